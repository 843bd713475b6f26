"""ctypes binding of libdust_gpu.so (the C ABI declared in include/dust_gpu.h).

The library is loaded from this package directory only.  There is no fallback:
if the shared object is missing or fails to load, every entry point raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdust_gpu.so")

DUST_OK = 0
DUST_E_ARG = -1
DUST_E_HIP = -2
DUST_E_OUTSIDE = -3
DUST_E_CAPACITY = -4
DUST_E_NOMEM = -5
DUST_E_UNSUPPORTED = -6

FMM_FP64 = 0
FMM_TUNED = 1
KERNEL_ROSENHEAD = 0
KERNEL_GAUSSIAN = 1

LU_FP64 = 0
LU_FP32 = 1
LU_FP16S = 2

STAGES = ("sort", "p2m", "m2m", "m2l", "l2l", "p2p", "l2p")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class FmmParams(C.Structure):
    _fields_ = [
        ("box_length", C.c_double),
        ("nbox", C.c_int * 3),
        ("origin", C.c_double * 3),
        ("nlevels", C.c_int),
        ("min_part", C.c_int),
        ("degree", C.c_int),
        ("delta", C.c_double),
        ("use_vs", C.c_int),
        ("use_vd", C.c_int),
        ("use_tv", C.c_int),
        ("nu_inf", C.c_double),
        ("u_inf", C.c_double * 3),
        ("precision", C.c_int),
        ("kernel", C.c_int),
        ("use_divfilt", C.c_int),
        ("reserved", C.c_int * 5),
    ]


class WindParams(C.Structure):
    _fields_ = [
        ("u_inf", C.c_double * 3),
        ("gust_type", C.c_int),
        ("reserved0", C.c_int),
        ("origin", C.c_double * 3),
        ("front_direction", C.c_double * 3),
        ("front_speed", C.c_double),
        ("u_des", C.c_double),
        ("perturb_direction", C.c_double * 3),
        ("gradient", C.c_double),
        ("start_time", C.c_double),
        ("reserved", C.c_int * 4),
    ]


GUST_NONE, GUST_ACM, GUST_LINEAR = 0, 1, 2


class AicParams(C.Structure):
    _fields_ = [
        ("ff_ratio_doublet", C.c_double),
        ("rankine_rad", C.c_double),
        ("cutoff_rad", C.c_double),
        ("reserved", C.c_int * 4),
    ]


ELEM_SURFPAN = 0
ELEM_RING = 1
ELEM_LINE = 2
ELEM_FP64 = 0
ELEM_TUNED = 1


class ElemGroup(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("n", C.c_int),
        ("n_ver", _vp),
        ("ver", _vp),
        ("mag", _vp),
        ("sigma", _vp),
        ("stretch", C.c_int),
        ("reserved", C.c_int),
    ]


class ElemParams(C.Structure):
    _fields_ = [
        ("ff_ratio_doublet", C.c_double),
        ("rankine_rad", C.c_double),
        ("cutoff_rad", C.c_double),
        ("vortex_rad", C.c_double),
        ("precision", C.c_int),
        ("reserved", C.c_int * 3),
    ]


class WakeParams(C.Structure):
    _fields_ = [
        ("box_min", C.c_double * 3),
        ("box_max", C.c_double * 3),
        ("dt", C.c_double),
        ("filt_eta", C.c_double),
        ("use_vs", C.c_int),
        ("use_vd", C.c_int),
        ("use_divfilt", C.c_int),
        ("reserved", C.c_int * 5),
    ]


class PaParams(C.Structure):
    _fields_ = [
        ("dt", C.c_double),
        ("dt_wake", C.c_double),
        ("u_ref", C.c_double),
        ("rad_mult", C.c_double),
        ("elrad_mult", C.c_double),
        ("reserved", C.c_int * 4),
    ]


class PanwakeParams(C.Structure):
    _fields_ = [
        ("n_stripes", C.c_int), ("n_points", C.c_int), ("nmax_pan", C.c_int),
        ("start_points", _vp), ("neigh", _vp), ("neigh_o", _vp), ("gen_elems", _vp),
        ("dt", C.c_double), ("ndt_update_wake", C.c_int), ("min_vel_at_te", C.c_double),
        ("join_te", C.c_int), ("join_te_factor", C.c_double),
        ("refine_wake", C.c_int), ("k_refine", C.c_int),
        ("k_vortex_rad", C.c_double), ("vortex_rad", C.c_double),
        ("box_min", C.c_double * 3), ("box_max", C.c_double * 3),
        ("reserved", C.c_int * 4),
    ]


ROW_SURFPAN = 0
ROW_VORTLATT = 1


class LinsysDesc(C.Structure):
    _fields_ = [
        ("rank", C.c_int), ("nstatic", C.c_int),
        ("kind", _vp), ("n_ver", _vp), ("ver", _vp), ("ub", _vp), ("uvort", _vp),
        ("bern", _vp),
        ("u_inf", C.c_double * 3),
        ("n_sp", C.c_int), ("nstatic_sp", C.c_int), ("id_sp", _vp),
        ("n_wake", C.c_int), ("n_impl", C.c_int),
        ("wake_nver", _vp), ("wake_ver", _vp), ("wake_mag", _vp), ("impl_ind", _vp),
        ("n_expl", C.c_int), ("nstatic_expl", C.c_int),
        ("expl_nver", _vp), ("expl_ver", _vp), ("expl_mag", _vp),
        ("ff_ratio_doublet", C.c_double), ("rankine_rad", C.c_double),
        ("cutoff_rad", C.c_double),
    ]


# (name, restype, argtypes) of every exported function in include/dust_gpu.h
SIGNATURES = [
    ("dust_last_error", C.c_char_p, []),
    ("dust_version", C.c_char_p, []),
    ("dust_fmm_create", C.c_int, [C.POINTER(FmmParams), C.c_int, C.POINTER(_vp)]),
    ("dust_fmm_destroy", C.c_int, [_vp]),
    ("dust_fmm_set_particles", C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp]),
    ("dust_fmm_set_particles_dev", C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp]),
    ("dust_fmm_sort_particles", C.c_int, [_vp, _ip]),
    ("dust_fmm_set_redistribution", C.c_int,
     [_vp, C.c_int, C.c_int, C.c_double, C.c_int, _vp]),
    ("dust_fmm_redistributed", C.c_int, [_vp, _ip, _vp, _vp, _vp]),
    ("dust_fmm_redistributed_dev", C.c_int, [_vp, _vp, _vp, _vp]),
    ("dust_fmm_calculate_multipole", C.c_int, [_vp]),
    ("dust_fmm_apply_multipole", C.c_int, [_vp, _vp, _vp, _vp]),
    ("dust_fmm_apply_multipole_dev", C.c_int, [_vp]),
    ("dust_fmm_rotu", C.c_int, [_vp, _vp]),
    ("dust_fmm_set_wind", C.c_int, [_vp, C.POINTER(WindParams), C.c_double]),
    ("dust_variable_wind", C.c_int, [C.c_int, _vp, C.POINTER(WindParams), C.c_double, _vp]),
    ("dust_variable_wind_dev", C.c_int,
     [C.c_int, _vp, C.POINTER(WindParams), C.c_double, _vp, _vp]),
    ("dust_fmm_rotu_dev", C.c_int, [_vp, _vp]),
    ("dust_fmm_results_dev", C.c_int, [_vp, _vp, _vp, _vp]),
    ("dust_fmm_apply_multipole_panels", C.c_int, [_vp, C.c_int, _vp, _vp]),
    ("dust_fmm_apply_multipole_panels_dev", C.c_int, [_vp, C.c_int, _vp, _vp]),
    ("dust_fmm_pass_dev", C.c_int, [_vp, _ip]),
    ("dust_fmm_pass_batch_dev", C.c_int, [C.c_int, _vp, _ip]),
    ("dust_fmm_graph_reserve", C.c_int, [_vp, C.c_double]),
    ("dust_fmm_pass_graph_dev", C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("dust_fmm_graph_capture", C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("dust_fmm_graph_launch", C.c_int, [_vp]),
    ("dust_fmm_graph_status", C.c_int, [_vp, _ip]),
    ("dust_fmm_set_tuning", C.c_int, [_vp, C.c_char_p, C.c_double]),
    ("dust_fmm_get_tuning", C.c_int, [_vp, C.c_char_p, _dp]),
    ("dust_fmm_autotune", C.c_int, [_vp, C.c_int, _dp]),
    ("dust_lu_set_tuning", C.c_int, [C.c_char_p, C.c_int]),
    ("dust_lu_get_tuning", C.c_int, [C.c_char_p, _vp]),
    ("dust_fmm_stream", _vp, [_vp]),
    ("dust_fmm_synchronize", C.c_int, [_vp]),
    ("dust_fmm_leaves", C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    ("dust_fmm_cell_index", C.c_int, [_vp, _vp]),
    ("dust_fmm_work", C.c_int, [_vp, _dp, _dp]),
    ("dust_fmm_set_timing", C.c_int, [_vp, C.c_int]),
    ("dust_fmm_stage_times", C.c_int, [_vp, _vp, _vp]),
    ("dust_vortpart_direct", C.c_int,
     [C.c_int, _vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_double, _vp, _vp, _vp]),
    ("dust_vortpart_direct_k", C.c_int,
     [C.c_int, _vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_double, _vp, C.c_int, _vp, _vp]),
    ("dust_elem_apply", C.c_int,
     [C.c_int, C.POINTER(ElemGroup), C.POINTER(ElemParams), C.c_int, _vp, _vp, _vp, _vp]),
    ("dust_elem_apply_dev", C.c_int,
     [C.c_int, C.POINTER(ElemGroup), C.POINTER(ElemParams), C.c_int, _vp, _vp, _vp, _vp, _vp]),
    ("dust_vort_vel", C.c_int,
     [C.c_int, _vp, C.c_int, _vp, _vp, _vp, _vp, C.c_int, _vp, _vp, C.POINTER(ElemParams), _vp]),
    ("dust_vort_vel_dev", C.c_int,
     [C.c_int, _vp, C.c_int, _vp, _vp, _vp, _vp, C.c_int, _vp, _vp, C.POINTER(ElemParams), _vp,
      _vp]),
    ("dust_wake_update_dev", C.c_int,
     [C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(WakeParams), _vp, _vp]),
    ("dust_wake_live_dev", C.c_int, [C.c_int, _vp, _vp, _vp, _vp]),
    ("dust_wake_avoid_collision_dev", C.c_int,
     [C.c_int, _vp, _vp, _vp, _vp, C.c_int, C.POINTER(ElemGroup), _vp, C.POINTER(PaParams),
      _vp]),
    ("dust_wake_shed_dev", C.c_int,
     [C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int, _vp, _vp, _vp, _vp, C.POINTER(WakeParams),
      _vp, _vp]),
    ("dust_wake_gather_dev", C.c_int,
     [C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("dust_wake_scatter_dev", C.c_int, [C.c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("dust_aic_rhs_dev", C.c_int,
     [C.c_int, _vp, _vp, _vp, C.c_int, C.c_int, _vp, _vp, _vp, C.c_int, _vp, _vp]),
    ("dust_aic_surfpan", C.c_int,
     [C.c_int, _vp, _vp, C.POINTER(AicParams), _vp, C.c_int, _vp, C.c_int]),
    ("dust_aic_surfpan_dev", C.c_int,
     [C.c_int, _vp, _vp, C.POINTER(AicParams), _vp, C.c_int, _vp, C.c_int, _vp]),
    ("dust_aic_vortlatt", C.c_int,
     [C.c_int, _vp, _vp, C.POINTER(AicParams), _vp, C.c_int, _vp, C.c_int]),
    ("dust_aic_rhs", C.c_int,
     [C.c_int, _vp, _vp, _vp, C.c_int, C.c_int, _vp, _vp, _vp, C.c_int]),
    ("dust_dgetrf", C.c_int, [C.c_int, C.c_int, _vp, C.c_int, _vp]),
    ("dust_dgetrs", C.c_int, [C.c_int, C.c_int, _vp, C.c_int, _vp, _vp, C.c_int]),
    ("dust_dgetrf_dev", C.c_int, [C.c_int, C.c_int, _vp, C.c_int, _vp, _vp]),
    ("dust_dgetrs_dev", C.c_int, [C.c_int, C.c_int, _vp, C.c_int, _vp, _vp, C.c_int, _vp]),
    ("dust_dgetrs_noalloc_dev", C.c_int,
     [C.c_int, C.c_int, _vp, C.c_int, _vp, _vp, C.c_int, _vp]),
    ("dust_dgesv_mixed", C.c_int,
     [C.c_int, C.c_int, _vp, C.c_int, _vp, C.c_int, C.c_int, C.c_double, C.c_int, _ip, _dp]),
    ("dust_dgesv_mixed_dev", C.c_int,
     [C.c_int, C.c_int, _vp, C.c_int, _vp, C.c_int, C.c_int, C.c_double, C.c_int, _ip, _dp,
      _vp]),
    ("dust_dgesv_mixed_batch_dev", C.c_int,
     [C.c_int, _ip, _vp, _ip, _vp, _ip, C.c_int, C.c_int, C.c_double, C.c_int, _ip, _dp]),
    ("dust_solve_linsys", C.c_int, [C.c_int, C.c_int, _vp, C.c_int, _vp, _vp, _vp]),
    ("dust_initialize_linsys", C.c_int,
     [_vp, _vp, C.c_int, _vp, C.c_int, _vp, C.c_int, _vp, C.c_int, _vp, _vp]),
    ("dust_initialize_linsys_dev", C.c_int,
     [_vp, _vp, C.c_int, _vp, C.c_int, _vp, C.c_int, _vp, C.c_int, _vp, _vp, _vp]),
    ("dust_assemble_linsys", C.c_int,
     [_vp, _vp, C.c_int, _vp, C.c_int, _vp, C.c_int, _vp, _vp, C.c_int, _vp]),
    ("dust_assemble_linsys_dev", C.c_int,
     [_vp, _vp, C.c_int, _vp, C.c_int, _vp, C.c_int, _vp, _vp, C.c_int, _vp, _vp]),
    ("dust_solve_pressure_sys", C.c_int, [C.c_int, C.c_int, _vp, C.c_int, _vp, _vp, _vp]),
    ("dust_press_normvel_der", C.c_int,
     [C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_double, _vp]),
    ("dust_press_normvel_der_dev", C.c_int,
     [C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_double, _vp, _vp]),
    ("dust_pres_rhs_src", C.c_int, [C.c_int, _vp, C.c_int, _vp, _vp]),
    ("dust_pres_rhs_src_dev", C.c_int, [C.c_int, _vp, C.c_int, _vp, _vp, _vp]),
    ("dust_pres_rhs_rot", C.c_int,
     [C.c_int, _vp, C.c_int, _vp, _vp, _vp, _vp, _vp, C.c_int, _vp, _vp, _vp, C.c_int, _vp, _vp,
      _vp, _vp, _vp]),
    ("dust_pres_rhs_rot_dev", C.c_int,
     [C.c_int, _vp, C.c_int, _vp, _vp, _vp, _vp, _vp, C.c_int, _vp, _vp, _vp, C.c_int, _vp, _vp,
      _vp, _vp, _vp, _vp]),
    ("dust_test_gemm_sub_dev", C.c_int,
     [C.c_int, C.c_int, C.c_int, _vp, C.c_int, _vp, C.c_int, _vp, C.c_int]),
    ("dust_test_trsm_llnu_dev", C.c_int, [C.c_int, C.c_int, _vp, C.c_int, _vp, C.c_int]),
    ("dust_test_getrs_set_error_word", C.c_int, [C.c_uint]),
    ("dust_shutdown", C.c_int, []),
    ("dust_panwake_create", C.c_int, [_vp, _vp]),
    ("dust_panwake_destroy", C.c_int, [_vp]),
    ("dust_panwake_start_dev", C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp, _dp, _vp]),
    ("dust_panwake_panels", C.c_int, [_vp, _ip, _ip, _vp, _vp, _vp, _vp]),
    ("dust_panwake_points", C.c_int, [_vp, _ip, _vp]),
    ("dust_panwake_update_dev", C.c_int, [_vp, _vp, _vp, _vp]),
    ("dust_panwake_first_row_dev", C.c_int, [_vp, _vp, _vp]),
    ("dust_panwake_shed_dev", C.c_int,
     [_vp, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("dust_panwake_end_vortices", C.c_int, [_vp, _ip, _vp, _vp, _vp]),
    ("dust_panwake_state", C.c_int, [_vp, _ip, _ip, _ip]),
    ("dust_panwake_checkpoint", C.c_int, [_vp, _vp, C.POINTER(C.c_size_t)]),
    ("dust_panwake_restore", C.c_int, [_vp, _vp, C.c_size_t]),
    ("dust_panwake_load", C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp, _vp]),
    ("dust_panwake_get", C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
]

_lib = None


class DustError(RuntimeError):
    def __init__(self, code, where):
        msg = ""
        if _lib is not None:
            raw = _lib.dust_last_error()
            msg = raw.decode() if raw else ""
        super().__init__(f"{where} failed with status {code}: {msg}")
        self.code = code


def lib():
    """Load libdust_gpu.so (no fallback: raises if it is absent)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python dust_amd/build.py` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        # PyTorch ships its own HIP runtime under the same soname
        # (libamdhip64.so.7).  Whichever is loaded first serves the whole
        # process, and torch fails to initialise on a runtime other than its
        # own, so torch (when installed) is loaded before libdust_gpu.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(code, where):
    """Raise DustError for a negative status; return LAPACK-style info (>= 0)."""
    if code < 0:
        raise DustError(code, where)
    return code


def ptr(a):
    """Address of a numpy array or torch tensor (None stays NULL)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data
