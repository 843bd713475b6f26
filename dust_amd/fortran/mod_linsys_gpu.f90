! mod_linsys_gpu -- DUST's linear-system solves (src/linsys) on libdust_gpu,
! with the reference's names:
!   t_linsys                the fields of mod_linsys_vars.f90:70-149 that the
!                           solves read and write (same names and meaning)
!   solve_linsys(linsys)    mod_linsys.f90:352-483: with skip = .false. the
!                           moving rows and columns are eliminated against the
!                           static block that initialize_linsys factored
!                           (laswp / trsm / gemm), the moving block is
!                           factored and the pivots made global; then
!                           res = A^-1 b.  With skip = .true. only the solve.
!   solve_pressure_sys(linsys)   mod_pressure_equation.f90:251-300, the same
!                           on A_pres / b_pres / P_pres over the n_sp surface
!                           panels (nstatic_sp static).
!   factorize_static(linsys)     the static-block dgetrf of initialize_linsys
!                           (mod_linsys.f90:247-262), for callers that build
!                           the matrix themselves.
! Errors stop with the library's message, where the reference calls error().
module mod_linsys_gpu
  use, intrinsic :: iso_c_binding
  use mod_dust_gpu
  implicit none
  private
  public :: t_linsys, solve_linsys, solve_pressure_sys, factorize_static

  integer, parameter :: wp = c_double

  type :: t_linsys
    integer :: rank = 0
    real(wp), allocatable :: A(:,:), b(:), A_pres(:,:), b_pres(:)
    integer(c_int), allocatable :: P(:), P_pres(:)
    real(wp), allocatable :: res(:), res_pres(:)
    integer :: nstatic = 0, nmoving = 0
    integer :: nstatic_sp = 0, nmoving_sp = 0, n_sp = 0
    logical :: skip = .false.
  end type

contains

  subroutine check(ierr, what)
    integer(c_int), intent(in) :: ierr
    character(len=*), intent(in) :: what
    if (ierr /= DUST_OK) then
      print '(a,a,a,i0,a,a)', 'mod_linsys_gpu: ', what, ' failed (', ierr, '): ', dust_error_text()
      error stop 1
    end if
  end subroutine check

  ! the static blocks of A and (when allocated) A_pres, factored in place
  subroutine factorize_static(linsys)
    type(t_linsys), intent(inout) :: linsys
    integer(c_int) :: info
    if (linsys%nstatic > 0) then
      info = dust_dgetrf(int(linsys%nstatic, c_int), int(linsys%nstatic, c_int), linsys%A, &
                         int(linsys%rank, c_int), linsys%P)
      if (info /= 0) then
        print '(a,i0)', 'mod_linsys_gpu: error while factorizing the static part of the &
              &linear system, DGETRF code ', info
        error stop 1
      end if
    end if
    if (linsys%nstatic_sp > 0 .and. allocated(linsys%A_pres)) then
      info = dust_dgetrf(int(linsys%nstatic_sp, c_int), int(linsys%nstatic_sp, c_int), &
                         linsys%A_pres, int(linsys%n_sp, c_int), linsys%P_pres)
      if (info /= 0) then
        print '(a,i0)', 'mod_linsys_gpu: error while factorizing the static part of the &
              &linear system for pressure, DGETRF code ', info
        error stop 1
      end if
    end if
  end subroutine factorize_static

  subroutine solve_linsys(linsys)
    type(t_linsys), intent(inout) :: linsys
    if (.not. linsys%skip) then
      call check(dust_solve_linsys(int(linsys%rank, c_int), int(linsys%nstatic, c_int), &
                 linsys%A, int(linsys%rank, c_int), linsys%P, linsys%b, linsys%res), &
                 'solve_linsys')
    else
      linsys%res = linsys%b
      call check(dust_dgetrs(int(linsys%rank, c_int), 1_c_int, linsys%A, &
                 int(linsys%rank, c_int), linsys%P, linsys%res, int(linsys%rank, c_int)), &
                 'solve_linsys')
    end if
  end subroutine solve_linsys

  subroutine solve_pressure_sys(linsys)
    type(t_linsys), intent(inout) :: linsys
    call check(dust_solve_pressure_sys(int(linsys%n_sp, c_int), int(linsys%nstatic_sp, c_int), &
               linsys%A_pres, int(linsys%n_sp, c_int), linsys%P_pres, linsys%b_pres, &
               linsys%res_pres), 'solve_pressure_sys')
  end subroutine solve_pressure_sys

end module mod_linsys_gpu
