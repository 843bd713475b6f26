! Test driver: a Fortran caller of libdust_gpu through mod_dust_gpu, the way
! DUST's mod_octree / mod_linsys would call it.  Reads a stream file with a
! particle set and a dense system, runs initialize_octree -> sort_particles ->
! calculate_multipole -> apply_multipole -> destroy_octree and dgetrf +
! dgetrs, and writes the results to a second stream file.
!   in : int32 n, nlevels, min_part, degree, nlu; real64 box_length,
!        origin(3), delta; cen(3,n), dir(3,n), mag(n), r_vortex(n);
!        A(nlu,nlu), b(nlu)
!   out: int32 nleaves, info; real64 vel(3,n), stretch(3,n), x(nlu)
program dust_gpu_driver
  use, intrinsic :: iso_c_binding
  use mod_dust_gpu
  implicit none
  character(len=1024) :: fin, fout
  integer(c_int) :: n, nlevels, min_part, degree, nlu, nleaves, info
  real(c_double) :: box_length, origin(3), delta
  real(c_double), allocatable :: cen(:,:), dir(:,:), mag(:), rv(:), vel(:,:), str(:,:), tv(:)
  real(c_double), allocatable :: A(:,:), b(:)
  integer(c_int), allocatable :: ipiv(:)
  type(t_dust_fmm_params) :: prm
  type(c_ptr) :: h
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) n, nlevels, min_part, degree, nlu
  read(u) box_length, origin, delta
  allocate(cen(3,n), dir(3,n), mag(n), rv(n), vel(3,n), str(3,n), tv(n))
  read(u) cen, dir, mag, rv
  allocate(A(nlu,nlu), b(nlu), ipiv(nlu))
  read(u) A, b
  close(u)

  prm%box_length = box_length
  prm%nbox = 1
  prm%origin = origin
  prm%nlevels = nlevels
  prm%min_part = min_part
  prm%degree = degree
  prm%delta = delta
  prm%use_vs = 1
  prm%use_vd = 0
  prm%use_tv = 0
  prm%nu_inf = 0.0_c_double
  prm%u_inf = 0.0_c_double
  prm%precision = DUST_FMM_FP64
  prm%reserved = 0
  call check(dust_fmm_create(prm, n, h), 'initialize_octree')
  call check(dust_fmm_set_particles(h, n, cen, dir, mag, rv), 'set_particles')
  call check(dust_fmm_sort_particles(h, nleaves), 'sort_particles')
  call check(dust_fmm_calculate_multipole(h), 'calculate_multipole')
  call check(dust_fmm_apply_multipole(h, vel, str, tv), 'apply_multipole')
  call check(dust_fmm_destroy(h), 'destroy_octree')

  info = dust_dgetrf(nlu, nlu, A, nlu, ipiv)
  if (info < 0) print '(a,i0,a,a)', 'dust_gpu: dgetrf failed (', info, '): ', dust_error_text()
  if (info == 0) call check(dust_dgetrs(nlu, 1, A, nlu, ipiv, b, nlu), 'dgetrs')

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) nleaves, info
  write(u) vel, str, b
  close(u)

contains

  subroutine check(ierr, what)
    integer(c_int), intent(in) :: ierr
    character(len=*), intent(in) :: what
    if (ierr /= DUST_OK) then
      print '(a,a,a,i0,a,a)', 'dust_gpu: ', what, ' failed (', ierr, '): ', dust_error_text()
      stop 1
    end if
  end subroutine check
end program dust_gpu_driver
