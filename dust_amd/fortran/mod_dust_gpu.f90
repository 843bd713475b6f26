! mod_dust_gpu -- ISO_C_BINDING interface of libdust_gpu.so (include/dust_gpu.h)
! for DUST's Fortran solver: the src/fmm entry points (initialize_octree,
! sort_particles, calculate_multipole, apply_multipole, destroy_octree of
! src/fmm/mod_octree.f90:251,703,1002,1139,694) and the src/linsys dense calls
! (dgetrf/dgetrs of src/linsys/mod_linsys.f90:236,435,473 and the block path of
! solve_linsys, :352-483).  Every function returns the library's int status.
module mod_dust_gpu
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: DUST_OK = 0, DUST_E_ARG = -1, DUST_E_HIP = -2, &
       DUST_E_OUTSIDE = -3, DUST_E_CAPACITY = -4, DUST_E_NOMEM = -5
  integer(c_int), parameter :: DUST_FMM_FP64 = 0, DUST_FMM_TUNED = 1
  integer(c_int), parameter :: DUST_KERNEL_ROSENHEAD = 0, DUST_KERNEL_GAUSSIAN = 1
  integer(c_int), parameter :: DUST_LU_FP64 = 0, DUST_LU_FP32 = 1, DUST_LU_FP16S = 2

  ! dust_fmm_params (include/dust_gpu.h), same field order
  type, bind(C) :: t_dust_fmm_params
    real(c_double)    :: box_length
    integer(c_int)    :: nbox(3)
    real(c_double)    :: origin(3)
    integer(c_int)    :: nlevels, min_part, degree
    real(c_double)    :: delta
    integer(c_int)    :: use_vs, use_vd, use_tv
    real(c_double)    :: nu_inf
    real(c_double)    :: u_inf(3)
    integer(c_int)    :: precision
    integer(c_int)    :: kernel = 0   ! DUST_KERNEL_ROSENHEAD
    integer(c_int)    :: use_divfilt = 0
    integer(c_int)    :: reserved(5)
  end type

  type, bind(C) :: t_dust_aic_params
    real(c_double) :: ff_ratio_doublet, rankine_rad, cutoff_rad
    integer(c_int) :: reserved(4)
  end type

  ! dust_elem_group / dust_elem_params (include/dust_gpu.h), same field order
  integer(c_int), parameter :: DUST_ELEM_SURFPAN = 0, DUST_ELEM_RING = 1, DUST_ELEM_LINE = 2
  integer(c_int), parameter :: DUST_ELEM_FP64 = 0, DUST_ELEM_TUNED = 1
  type, bind(C) :: t_dust_elem_group
    integer(c_int) :: kind, n
    type(c_ptr)    :: n_ver, ver, mag, sigma
    integer(c_int) :: stretch, reserved
  end type

  type, bind(C) :: t_dust_elem_params
    real(c_double) :: ff_ratio_doublet, rankine_rad, cutoff_rad, vortex_rad
    integer(c_int) :: precision, reserved(3)
  end type

  interface
    integer(c_int) function dust_fmm_create(prm, capacity, h) bind(C)
      import; type(t_dust_fmm_params), intent(in) :: prm
      integer(c_int), value :: capacity; type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function dust_fmm_destroy(h) bind(C)
      import; type(c_ptr), value :: h
    end function
    integer(c_int) function dust_fmm_set_particles(h, n, cen, dir, mag, rv) bind(C)
      import; type(c_ptr), value :: h; integer(c_int), value :: n
      real(c_double), intent(in) :: cen(3,*), dir(3,*), mag(*), rv(*)
    end function
    integer(c_int) function dust_fmm_sort_particles(h, nleaves) bind(C)
      import; type(c_ptr), value :: h; integer(c_int), intent(out) :: nleaves
    end function
    integer(c_int) function dust_fmm_set_redistribution(h, on, lvl_solid, ratio, m, elem_cen) &
        bind(C)
      import; type(c_ptr), value :: h; integer(c_int), value :: on, lvl_solid, m
      real(c_double), value :: ratio; real(c_double), intent(in) :: elem_cen(3,*)
    end function
    integer(c_int) function dust_fmm_redistributed(h, n_kept, keep, dir, mag) bind(C)
      import; type(c_ptr), value :: h; integer(c_int), intent(out) :: n_kept
      integer(c_int8_t), intent(out) :: keep(*)
      real(c_double), intent(inout) :: dir(3,*), mag(*)
    end function
    integer(c_int) function dust_fmm_rotu(h, rotu) bind(C)
      import; type(c_ptr), value :: h; real(c_double), intent(out) :: rotu(3,*)
    end function
    integer(c_int) function dust_fmm_calculate_multipole(h) bind(C)
      import; type(c_ptr), value :: h
    end function
    integer(c_int) function dust_fmm_apply_multipole(h, vel, stretch, turbvisc) bind(C)
      import; type(c_ptr), value :: h
      real(c_double), intent(out) :: vel(3,*), stretch(3,*), turbvisc(*)
    end function
    integer(c_int) function dust_aic_surfpan(n, n_ver, ver, prm, A, lda, B, ldb) bind(C)
      import; integer(c_int), value :: n, lda, ldb
      integer(c_int), intent(in) :: n_ver(*); real(c_double), intent(in) :: ver(3,4,*)
      type(t_dust_aic_params), intent(in) :: prm
      real(c_double), intent(out) :: A(lda,*), B(ldb,*)
    end function
    integer(c_int) function dust_dgetrf(m, n, A, lda, ipiv) bind(C)
      import; integer(c_int), value :: m, n, lda
      real(c_double), intent(inout) :: A(lda,*); integer(c_int), intent(out) :: ipiv(*)
    end function
    integer(c_int) function dust_dgetrs(n, nrhs, A, lda, ipiv, B, ldb) bind(C)
      import; integer(c_int), value :: n, nrhs, lda, ldb
      real(c_double), intent(in) :: A(lda,*); integer(c_int), intent(in) :: ipiv(*)
      real(c_double), intent(inout) :: B(ldb,*)
    end function
    integer(c_int) function dust_solve_linsys(rank, nstatic, A, lda, ipiv, b, res) bind(C)
      import; integer(c_int), value :: rank, nstatic, lda
      real(c_double), intent(inout) :: A(lda,*); integer(c_int), intent(inout) :: ipiv(*)
      real(c_double), intent(in) :: b(*); real(c_double), intent(out) :: res(*)
    end function
    integer(c_int) function dust_solve_pressure_sys(n_sp, nstatic_sp, A_pres, lda, P_pres, &
                                                    b_pres, res_pres) bind(C)
      import; integer(c_int), value :: n_sp, nstatic_sp, lda
      real(c_double), intent(inout) :: A_pres(lda,*); integer(c_int), intent(inout) :: P_pres(*)
      real(c_double), intent(in) :: b_pres(*); real(c_double), intent(out) :: res_pres(*)
    end function
    integer(c_int) function dust_pres_rhs_src(n_sp, B, ldb, src, b_pres) bind(C)
      import; integer(c_int), value :: n_sp, ldb
      real(c_double), intent(in) :: B(ldb,*), src(*); real(c_double), intent(out) :: b_pres(*)
    end function
    integer(c_int) function dust_pres_rhs_rot(m, cen, np, pcen, pdir, pmag, prv, pvel, nl, &
                                              lver, lmag, lvel, nw, wnver, wver, wmag, wvel, &
                                              b_pres) bind(C)
      import; integer(c_int), value :: m, np, nl, nw
      real(c_double), intent(in) :: cen(3,*)
      type(c_ptr), value :: pcen, pdir, pmag, prv, pvel, lver, lmag, lvel, wnver, wver, wmag, wvel
      real(c_double), intent(inout) :: b_pres(*)
    end function
    integer(c_int) function dust_dgesv_mixed(n, nrhs, A, lda, B, ldb, prec, tol, max_iter, &
                                             iters, resid) bind(C)
      import; integer(c_int), value :: n, nrhs, lda, ldb, prec, max_iter
      real(c_double), intent(in) :: A(lda,*); real(c_double), intent(inout) :: B(ldb,*)
      real(c_double), value :: tol
      integer(c_int), intent(out) :: iters; real(c_double), intent(out) :: resid
    end function
    type(c_ptr) function dust_last_error() bind(C)
      import
    end function
    integer(c_int) function dust_aic_vortlatt(n, n_ver, ver, prm, A, lda, B, ldb) bind(C)
      import; integer(c_int), value :: n, lda, ldb
      integer(c_int), intent(in) :: n_ver(*); real(c_double), intent(in) :: ver(3,4,*)
      type(t_dust_aic_params), intent(in) :: prm
      real(c_double), intent(out) :: A(lda,*), B(ldb,*)
    end function
    integer(c_int) function dust_fmm_work(h, p2p_pairs, m2l_pairs) bind(C)
      import; type(c_ptr), value :: h; real(c_double), intent(out) :: p2p_pairs, m2l_pairs
    end function
    integer(c_int) function dust_elem_apply(ngroups, groups, prm, n, pos, alpha, vel, &
                                            stretch) bind(C)
      import; integer(c_int), value :: ngroups, n
      type(t_dust_elem_group), intent(in) :: groups(*)
      type(t_dust_elem_params), intent(in) :: prm
      real(c_double), intent(in) :: pos(3,*)
      type(c_ptr), value :: alpha, stretch
      real(c_double), intent(inout) :: vel(3,*)
    end function
    integer(c_int) function dust_vort_vel(m, cen, np, pcen, pdir, pmag, prv, nl, lver, lmag, &
                                          prm, uvort) bind(C)
      import; integer(c_int), value :: m, np, nl
      real(c_double), intent(in) :: cen(3,*)
      type(c_ptr), value :: pcen, pdir, pmag, prv, lver, lmag
      type(t_dust_elem_params), intent(in) :: prm
      real(c_double), intent(out) :: uvort(3,*)
    end function
    integer(c_int) function dust_fmm_apply_multipole_panels(h, m, cen, uvort) bind(C)
      import; type(c_ptr), value :: h; integer(c_int), value :: m
      real(c_double), intent(in) :: cen(3,*); real(c_double), intent(inout) :: uvort(3,*)
    end function
  end interface

contains

  ! Text of the library's last error as a Fortran string.
  function dust_error_text() result(s)
    character(len=:), allocatable :: s
    character(kind=c_char), pointer :: p(:)
    type(c_ptr) :: cp
    integer :: k
    cp = dust_last_error()
    s = ''
    if (.not. c_associated(cp)) return
    call c_f_pointer(cp, p, [4096])
    do k = 1, 4096
      if (p(k) == c_null_char) exit
      s = s // p(k)
    end do
  end function
end module mod_dust_gpu
