! Test driver of the reference-named Fortran modules (mod_octree_gpu,
! mod_linsys_gpu): the calls of DUST's prepare_wake / update_wake
! (src/aeroel/mod_wake.f90:797-801, 1117) and of the time step's solve
! (src/dust.f90 -> solve_linsys), in that order, on data read from a stream
! file; the results go to a second stream file.
!   in : int32 nmax, n_prt, nlevels, min_part, degree, use_pr, lvl_solid,
!        ne, nw, nr, nl, rank, nstatic;
!        real64 box_length, origin(3), delta, ratio, u_inf(3), ff_ratio,
!        rankine, cutoff, vortex_rad;
!        slots: cen(3,nmax), dir(3,nmax), mag(nmax), r_vortex(nmax),
!        int32 free(nmax);
!        elem: ver(3,4,ne), int32 n_ver(ne), mag(ne), sigma(ne), cen(3,ne);
!        wpan: ver(3,4,nw), mag(nw); wrin: ver(3,4,nr), mag(nr);
!        wvort: ver(3,2,nl), mag(nl);
!        linsys: A(rank,rank), b(rank), b2(rank)
!   out: int32 n_prt, nleaves, free(nmax); real64 dir(3,nmax), mag(nmax),
!        vel(3,nmax), stretch(3,nmax), uvort(3,ne), res(rank), res2(rank)
program dust_shim_driver
  use, intrinsic :: iso_c_binding
  use mod_dust_gpu, only: DUST_ELEM_SURFPAN, DUST_ELEM_RING, DUST_ELEM_LINE
  use mod_octree_gpu
  use mod_linsys_gpu
  implicit none
  integer, parameter :: wp = c_double
  character(len=1024) :: fin, fout
  integer(c_int) :: hdr(13)
  real(wp) :: rh(13)
  integer :: nmax, n_prt, nlevels, min_part, degree, ne, nw, nr, nl, rank, nstatic, u, k
  integer(c_int), allocatable :: freei(:)
  real(wp), allocatable :: b2(:), res1(:)
  type(t_vortparts) :: wparts
  type(t_elems) :: elem, wpan, wrin, wvort
  type(t_octree) :: octree
  type(t_linsys) :: linsys

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) hdr
  nmax = hdr(1); n_prt = hdr(2); nlevels = hdr(3); min_part = hdr(4); degree = hdr(5)
  ne = hdr(8); nw = hdr(9); nr = hdr(10); nl = hdr(11); rank = hdr(12); nstatic = hdr(13)
  read(u) rh
  allocate(wparts%cen(3,nmax), wparts%dir(3,nmax), wparts%mag(nmax), wparts%r_Vortex(nmax), &
           wparts%vel(3,nmax), wparts%stretch(3,nmax), wparts%turbvisc(nmax), &
           wparts%free(nmax), freei(nmax))
  read(u) wparts%cen, wparts%dir, wparts%mag, wparts%r_Vortex, freei
  wparts%free = freei /= 0
  wparts%vel = 0.0_wp; wparts%stretch = 0.0_wp; wparts%turbvisc = 0.0_wp
  call read_list(elem, DUST_ELEM_SURFPAN, ne, 4, .true.)
  call read_list(wpan, DUST_ELEM_RING, nw, 4, .false.)
  call read_list(wrin, DUST_ELEM_RING, nr, 4, .false.)
  call read_list(wvort, DUST_ELEM_LINE, nl, 2, .false.)
  linsys%rank = rank
  linsys%nstatic = nstatic
  linsys%nmoving = rank - nstatic
  allocate(linsys%A(rank,rank), linsys%b(rank), linsys%P(rank), linsys%res(rank), b2(rank), &
           res1(rank))
  read(u) linsys%A, linsys%b, b2
  close(u)

  ! sim_param
  gpu_sim_param%use_vs = .true.
  gpu_sim_param%use_vd = .false.
  gpu_sim_param%vs_elems = .false.
  gpu_sim_param%use_pr = hdr(6) /= 0
  gpu_sim_param%lvl_solid = hdr(7)
  gpu_sim_param%part_redist_ratio = rh(6)
  gpu_sim_param%u_inf = rh(7:9)
  gpu_sim_param%FarFieldRatioDoublet = rh(10)
  gpu_sim_param%RankineRad = rh(11)
  gpu_sim_param%CutoffRad = rh(12)
  gpu_sim_param%VortexRad = rh(13)

  ! prepare_wake (mod_wake.f90:797-801), then update_wake's apply_multipole
  call initialize_octree(rh(1), [1, 1, 1], rh(2:4), nlevels, min_part, degree, rh(5), octree)
  call sort_particles(wparts, n_prt, elem, octree)
  call calculate_multipole(wparts, octree)
  if (allocated(elem%uvort)) elem%uvort = 0.0_wp
  call apply_multipole_panels(octree, elem)
  call apply_multipole(wparts, octree, elem, wpan, wrin, wvort)

  ! the solve of a step, then a solve that reuses the factors (skip)
  call factorize_static(linsys)
  call solve_linsys(linsys)
  res1 = linsys%res
  linsys%skip = .true.
  linsys%b = b2
  call solve_linsys(linsys)

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) int(n_prt, c_int), int(octree%nleaves, c_int)
  write(u) (merge(1_c_int, 0_c_int, wparts%free(k)), k = 1, nmax)
  write(u) wparts%dir, wparts%mag, wparts%vel, wparts%stretch, elem%uvort, res1, linsys%res
  close(u)
  call destroy_octree(octree)

contains

  subroutine read_list(e, kind, n, nv, surf)
    type(t_elems), intent(inout) :: e
    integer, intent(in) :: kind, n, nv
    logical, intent(in) :: surf
    e%kind = kind
    e%n = n
    allocate(e%ver(3, nv, max(n, 1)), e%mag(max(n, 1)))
    if (nv == 4) allocate(e%n_ver(max(n, 1)))
    if (n == 0) return
    read(u) e%ver(:, :, 1:n)
    if (surf) then
      read(u) e%n_ver(1:n)
      allocate(e%sigma(n), e%cen(3, n), e%uvort(3, n))
      read(u) e%mag(1:n), e%sigma, e%cen
    else
      if (nv == 4) e%n_ver = 4
      read(u) e%mag(1:n)
    end if
  end subroutine read_list
end program dust_shim_driver
