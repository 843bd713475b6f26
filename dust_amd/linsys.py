"""Host mirror of DUST's src/linsys dense kernels on top of libdust_gpu.

    aic_surfpan(ver, n_ver)     static AIC of surface panels: build_row_static_surfpan
                                with compute_pot_surfpan (src/aeroel/mod_surfpan.f90:448-481,
                                579-601)  -> (A, B)
    aic_vortlatt(ver, n_ver)    static AIC of vortex rings: build_row_static_vortlatt with
                                compute_psi_vortlatt (src/aeroel/mod_vortlatt.f90:201-237,
                                365-391)  -> (A, B)
    rhs_static(...)             static RHS (mod_surfpan.f90:400-404)
    dgetrf / dgetrs             the LAPACK call sites of src/linsys/mod_linsys.f90 (236, 435,
                                473): same arguments, 1-based ipiv, LAPACK info
    solve_linsys(...)           the block static/moving path, mod_linsys.f90:352-483
    dgesv_mixed(...)            tuned mixed-precision LU + FP64 iterative refinement

Matrices are numpy float64 arrays in Fortran (column-major) order.  A LAPACK
info > 0 is returned, as the reference receives it before calling error()
(mod_linsys.f90:241-244); API misuse raises DustError.
"""
import ctypes as C

import numpy as np

from . import _lib as L


def _params(ff_ratio_doublet=10.0, rankine_rad=0.1, cutoff_rad=0.001):
    p = L.AicParams()
    p.ff_ratio_doublet = float(ff_ratio_doublet)
    p.rankine_rad = float(rankine_rad)
    p.cutoff_rad = float(cutoff_rad)
    return p


def _panels(ver, n_ver):
    n = len(n_ver)
    v = np.zeros((n, 4, 3))
    ver = np.asarray(ver, dtype=np.float64)
    v[:, : ver.shape[1], :] = ver[:, :4, :]
    return np.ascontiguousarray(v), np.ascontiguousarray(n_ver, dtype=np.int32)


def aic_surfpan(ver, n_ver, with_source=True, **prm):
    v, nv = _panels(ver, n_ver)
    n = len(nv)
    A = np.zeros((n, n), order="F")
    B = np.zeros((n, n), order="F") if with_source else None
    L.check(L.lib().dust_aic_surfpan(n, L.ptr(nv), L.ptr(v), C.byref(_params(**prm)), L.ptr(A), n,
                                     L.ptr(B), n), "aic_surfpan")
    return A, B


def aic_vortlatt(ver, n_ver, **prm):
    v, nv = _panels(ver, n_ver)
    n = len(nv)
    A = np.zeros((n, n), order="F")
    B = np.zeros((n, n), order="F")
    L.check(L.lib().dust_aic_vortlatt(n, L.ptr(nv), L.ptr(v), C.byref(_params(**prm)), L.ptr(A), n,
                                      L.ptr(B), n), "aic_vortlatt")
    return A, B


def rhs_static(ver, n_ver, B, u_inf, uvort=None):
    """b(:, r) = B n.(-u_inf(:, r) - uvort) for each free-stream vector."""
    v, nv = _panels(ver, n_ver)
    n = len(nv)
    B = np.asfortranarray(B, dtype=np.float64)
    u = np.ascontiguousarray(np.atleast_2d(u_inf), dtype=np.float64)
    nrhs = u.shape[0]
    b = np.zeros((n, nrhs), order="F")
    uv = None if uvort is None else np.ascontiguousarray(uvort, dtype=np.float64)
    L.check(L.lib().dust_aic_rhs(n, L.ptr(nv), L.ptr(v), L.ptr(B), n, nrhs, L.ptr(u), L.ptr(uv),
                                 L.ptr(b), n), "rhs_static")
    return b


def dgetrf(A):
    """In-place LU of a Fortran-ordered float64 array.  Returns (ipiv, info)."""
    if not (A.flags.f_contiguous and A.dtype == np.float64):
        raise ValueError("A must be a Fortran-ordered float64 array")
    m, n = A.shape
    ipiv = np.zeros(min(m, n), np.int32)
    info = L.check(L.lib().dust_dgetrf(m, n, L.ptr(A), m, L.ptr(ipiv)), "dgetrf")
    return ipiv, info


def dgetrs(LU, ipiv, B):
    """In-place solve with the factors of dgetrf; B is (n,) or (n, nrhs)."""
    n = LU.shape[0]
    Bv = B.reshape(n, -1, order="F")
    if not (B.flags.f_contiguous and B.dtype == np.float64):
        raise ValueError("B must be a Fortran-ordered float64 array")
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    L.check(L.lib().dust_dgetrs(n, Bv.shape[1], L.ptr(LU), n, L.ptr(ipiv), L.ptr(B), n), "dgetrs")
    return B


def solve_linsys(A, ipiv, nstatic, b):
    """solve_linsys (mod_linsys.f90:352-483).  A(1:nstatic,1:nstatic) holds the
    static factors and ipiv(1:nstatic) their pivots; A and ipiv are updated in
    place as the reference leaves them.  Returns (res, info)."""
    rank = A.shape[0]
    res = np.zeros(rank)
    b = np.ascontiguousarray(b, dtype=np.float64)
    info = L.check(L.lib().dust_solve_linsys(rank, nstatic, L.ptr(A), rank, L.ptr(ipiv), L.ptr(b),
                                             L.ptr(res)), "solve_linsys")
    return res, info


def dgesv_mixed(A, B, prec="fp16s", tol=1e-12, max_iter=30):
    """Tuned solve of A X = B: FP32 or row-scaled FP16 LU on the matrix cores
    (FP64 diagonal blocks, threshold-checked block pivoting), FP64 residuals
    and iterative refinement.  A is not modified; B (Fortran order, (n,) or
    (n, nrhs)) is overwritten with X.  Returns (X, iters, backward_error);
    iters == -1 means the block pivots failed the check or refinement
    stalled and X came from the FP64 partial-pivoting path."""
    n = A.shape[0]
    if not (A.flags.f_contiguous and A.dtype == np.float64):
        raise ValueError("A must be a Fortran-ordered float64 array")
    if not (B.flags.f_contiguous and B.dtype == np.float64):
        raise ValueError("B must be a Fortran-ordered float64 array")
    Bv = B.reshape(n, -1, order="F")
    it = C.c_int()
    res = C.c_double()
    code = {"fp64": L.LU_FP64, "fp32": L.LU_FP32, "fp16s": L.LU_FP16S}[prec]
    info = L.check(L.lib().dust_dgesv_mixed(n, Bv.shape[1], L.ptr(A), n, L.ptr(B), n, code,
                                            float(tol), int(max_iter), C.byref(it), C.byref(res)),
                   "dgesv_mixed")
    if info > 0:
        raise np.linalg.LinAlgError(f"dgesv_mixed: info = {info}")
    return B, it.value, res.value


# ---- initialize_linsys / assemble_linsys with the wake ------------------------
ROW_SURFPAN = L.ROW_SURFPAN
ROW_VORTLATT = L.ROW_VORTLATT


class System:
    """Inputs of initialize_linsys / assemble_linsys (src/linsys/mod_linsys.f90:
    109-272, 279-344): the implicit elements (row kinds ROW_SURFPAN /
    ROW_VORTLATT, vertices, the nstatic static ones first), their body
    velocity ub and uvort, the Bernoulli source bern of the surface panels,
    the surface-panel list id_sp (0-based, the nstatic_sp static ones first),
    the wake panels (`wake`: dict ver, n_ver, mag, n_impl, impl_ind (n_impl, 2)
    1-based pan_gen_elems_id with 0 for a missing lower element) and the
    explicit elements (`expl`: dict ver, n_ver, mag, nstatic)."""

    def __init__(self, kind, ver, n_ver, nstatic, ub=None, uvort=None, bern=None,
                 u_inf=(0.0, 0.0, 0.0), id_sp=None, nstatic_sp=0, wake=None, expl=None,
                 ff_ratio=10.0, rankine=0.1, cutoff=0.001):
        self.ver, self.n_ver = _panels(ver, n_ver)
        self.rank = len(self.n_ver)
        self.kind = np.ascontiguousarray(kind, dtype=np.int32)
        self.nstatic = int(nstatic)
        f = lambda a, sh: None if a is None else np.ascontiguousarray(
            np.asarray(a, np.float64).reshape(sh))
        self.ub = f(ub, (self.rank, 3))
        self.uvort = f(uvort, (self.rank, 3))
        self.bern = f(bern, (self.rank,))
        self.u_inf = tuple(float(x) for x in u_inf)
        self.id_sp = np.zeros(0, np.int32) if id_sp is None else np.ascontiguousarray(
            id_sp, dtype=np.int32)
        self.n_sp = len(self.id_sp)
        self.nstatic_sp = int(nstatic_sp)
        w = wake or {}
        if w:
            self.wake_ver, self.wake_nver = _panels(w["ver"], w["n_ver"])
            self.wake_mag = f(w.get("mag", np.zeros(len(self.wake_nver))), (-1,))
            self.n_impl = int(w.get("n_impl", 0))
            self.impl_ind = np.ascontiguousarray(
                np.asarray(w.get("impl_ind", np.zeros((0, 2)))).reshape(-1, 2), dtype=np.int32)
        else:
            self.wake_ver, self.wake_nver, self.wake_mag = None, None, None
            self.n_impl, self.impl_ind = 0, np.zeros((0, 2), np.int32)
        self.n_wake = 0 if self.wake_nver is None else len(self.wake_nver)
        e = expl or {}
        if e:
            self.expl_ver, self.expl_nver = _panels(e["ver"], e["n_ver"])
            self.expl_mag = f(e["mag"], (-1,))
            self.nstatic_expl = int(e.get("nstatic", 0))
        else:
            self.expl_ver, self.expl_nver, self.expl_mag = None, None, None
            self.nstatic_expl = 0
        self.n_expl = 0 if self.expl_nver is None else len(self.expl_nver)
        self.ff_ratio, self.rankine, self.cutoff = float(ff_ratio), float(rankine), float(cutoff)

    ARRAYS = ("kind", "n_ver", "ver", "ub", "uvort", "bern", "id_sp", "wake_nver", "wake_ver",
              "wake_mag", "impl_ind", "expl_nver", "expl_ver", "expl_mag")

    def desc(self, ptrs=None):
        """LinsysDesc over the host arrays (or over `ptrs`: name -> address)."""
        d = L.LinsysDesc()
        d.rank, d.nstatic = self.rank, self.nstatic
        for k in self.ARRAYS:
            a = getattr(self, k)
            setattr(d, k, (ptrs or {}).get(k) if ptrs is not None else
                    (None if a is None or a.size == 0 else a.ctypes.data))
        d.u_inf = (C.c_double * 3)(*self.u_inf)
        d.n_sp, d.nstatic_sp = self.n_sp, self.nstatic_sp
        d.n_wake, d.n_impl = self.n_wake, self.n_impl
        d.n_expl, d.nstatic_expl = self.n_expl, self.nstatic_expl
        d.ff_ratio_doublet, d.rankine_rad, d.cutoff_rad = self.ff_ratio, self.rankine, self.cutoff
        return d


def initialize_linsys(S):
    """Returns dict(A, B, L_static, A_pres, ipiv, ipiv_pres, info); A and
    A_pres hold the LU factors of their static blocks."""
    n, nst = S.rank, S.nstatic
    st = dict(A=np.zeros((n, n), order="F"), B=np.zeros((n, n), order="F"),
              L_static=np.zeros((max(nst, 1), max(S.nstatic_expl, 1)), order="F"),
              A_pres=np.zeros((max(S.n_sp, 1), max(S.n_sp, 1)), order="F"),
              ipiv=np.zeros(max(nst, 1), np.int32), ipiv_pres=np.zeros(max(S.nstatic_sp, 1),
                                                                        np.int32))
    d = S.desc()
    st["info"] = L.check(L.lib().dust_initialize_linsys(
        C.byref(d), L.ptr(st["A"]), n, L.ptr(st["B"]), n, L.ptr(st["L_static"]),
        st["L_static"].shape[0], L.ptr(st["A_pres"]), st["A_pres"].shape[0], L.ptr(st["ipiv"]),
        L.ptr(st["ipiv_pres"])), "initialize_linsys")
    return st


def assemble_linsys(S, st, pressure=False):
    """One assembly into the state of initialize_linsys (A, B, A_pres updated
    in place).  Returns b, or (b, b_pres) with pressure=True."""
    n = S.rank
    b = np.zeros(n)
    bp = np.zeros(max(S.n_sp, 1)) if pressure else None
    d = S.desc()
    L.check(L.lib().dust_assemble_linsys(
        C.byref(d), L.ptr(st["A"]), n, L.ptr(st["B"]), n, L.ptr(st["L_static"]),
        st["L_static"].shape[0], L.ptr(b), L.ptr(st["A_pres"]) if S.n_sp else None,
        st["A_pres"].shape[0], L.ptr(bp)), "assemble_linsys")
    return (b, bp[:S.n_sp]) if pressure else b
