! mod_octree_gpu -- DUST's src/fmm interface (mod_octree.f90:78-79) on
! libdust_gpu, with the reference's procedure names, argument order and
! argument meaning:
!   initialize_octree(box_length, nbox, origin, nlevels, min_part, degree,
!                     delta, octree)                         mod_octree.f90:251-518
!   sort_particles(wparts, n_prt, elem, octree)                         703-996
!   calculate_multipole(part, octree)                                   1002-1133
!   apply_multipole(part, octree, elem, wpan, wrin, wvort)              1139-1429
!   apply_multipole_panels(octree, elem)                                1434-1510
!   destroy_octree(octree)                                              694
! The reference's polymorphic particle and element types have no
! C-interoperable layout; here they are the SoA containers t_vortparts (the
! wake's particle slots: wake_parts with their free flags) and t_elems (one
! element list: elem, wpan, wrin or wvort).  The sim_param switches the
! reference reads inside these procedures are the fields of gpu_sim_param,
! with the reference's defaults (mod_sim_param.f90:370-455).
! Errors stop with the library's message, where the reference calls error().
module mod_octree_gpu
  use, intrinsic :: iso_c_binding
  use mod_dust_gpu
  implicit none
  private
  public :: t_octree, t_vortparts, t_elems, t_gpu_sim_param, gpu_sim_param, &
            initialize_octree, sort_particles, calculate_multipole, apply_multipole, &
            apply_multipole_panels, destroy_octree

  integer, parameter :: wp = c_double

  ! the sim_param fields read by mod_octree and the element loops
  type :: t_gpu_sim_param
    logical  :: use_vs = .true.          ! vortstretch
    logical  :: use_vd = .true.          ! diffusion
    logical  :: use_tv = .false.         ! turbulent_viscosity
    logical  :: use_divfilt = .false.    ! divergence_filtering (rotu)
    real(wp) :: nu_inf = 0.0_wp          ! mu_inf / rho_inf of the case
    real(wp) :: u_inf(3) = 0.0_wp        ! uniform wind (variable_wind without gust)
    logical  :: vs_elems = .false.       ! vortstretch_from_elems
    real(wp) :: FarFieldRatioDoublet = 10.0_wp, RankineRad = 0.1_wp
    real(wp) :: CutoffRad = 0.001_wp, VortexRad = 0.1_wp
    logical  :: use_pr = .false.         ! particles_redistribution
    integer  :: lvl_solid = -1           ! octree_level_solid (-1: max(nlevels-2, 1))
    real(wp) :: part_redist_ratio = 3.0_wp
    logical  :: tuned = .false.          ! tuned form of the FMM and the element far fields
  end type
  type(t_gpu_sim_param), save :: gpu_sim_param

  ! t_octree: the device handle plus the packing of the last sort (slot of
  ! every particle passed to the device, in the reference's visiting order)
  type :: t_octree
    type(c_ptr) :: gpu = c_null_ptr
    type(t_dust_fmm_params) :: prm
    integer :: cap = 0
    integer :: nlevels = 0, lvl_solid = 1
    integer :: n = 0, nleaves = 0
    integer, allocatable :: slot(:)
    real(wp), allocatable :: cen(:,:), dir(:,:), mag(:), rv(:)
  end type

  ! the wake's particle slots (t_vortpart fields used by the FMM)
  type :: t_vortparts
    real(wp), allocatable :: cen(:,:), dir(:,:), mag(:), r_Vortex(:)
    real(wp), allocatable :: vel(:,:), stretch(:,:), turbvisc(:)
    real(wp), allocatable :: rotu(:,:)   ! with use_divfilt and use_vs
    logical, allocatable :: free(:)
  end type

  ! one element list: kind DUST_ELEM_SURFPAN / RING / LINE, vertices
  ! ver(3,4,n) (ver(3,2,n) for lines), n_ver(n) (polygons), intensities mag(n),
  ! source strengths sigma(n) (surface panels), centres cen(3,n) and the
  ! particle velocity at them uvort(3,n) (apply_multipole_panels)
  type :: t_elems
    integer :: kind = DUST_ELEM_SURFPAN
    integer :: n = 0
    real(wp), allocatable :: ver(:,:,:), mag(:), sigma(:), cen(:,:), uvort(:,:)
    integer(c_int), allocatable :: n_ver(:)
  end type

contains

  subroutine check(ierr, what)
    integer(c_int), intent(in) :: ierr
    character(len=*), intent(in) :: what
    if (ierr /= DUST_OK) then
      print '(a,a,a,i0,a,a)', 'mod_octree_gpu: ', what, ' failed (', ierr, '): ', dust_error_text()
      error stop 1
    end if
  end subroutine check

  subroutine initialize_octree(box_length, nbox, origin, nlevels, min_part, degree, delta, &
                               octree)
    real(wp), intent(in) :: box_length
    integer, intent(in) :: nbox(3)
    real(wp), intent(in) :: origin(3)
    integer, intent(in) :: nlevels, min_part, degree
    real(wp), intent(in) :: delta
    type(t_octree), intent(out) :: octree
    type(t_dust_fmm_params) :: prm
    integer, parameter :: capacity = 1024
    prm%box_length = box_length
    prm%nbox = int(nbox, c_int)
    prm%origin = origin
    prm%nlevels = int(nlevels, c_int)
    prm%min_part = int(min_part, c_int)
    prm%degree = int(degree, c_int)
    prm%delta = delta
    prm%use_vs = merge(1_c_int, 0_c_int, gpu_sim_param%use_vs)
    prm%use_vd = merge(1_c_int, 0_c_int, gpu_sim_param%use_vd)
    prm%use_tv = merge(1_c_int, 0_c_int, gpu_sim_param%use_tv)
    prm%nu_inf = gpu_sim_param%nu_inf
    prm%u_inf = gpu_sim_param%u_inf
    prm%precision = merge(DUST_FMM_TUNED, DUST_FMM_FP64, gpu_sim_param%tuned)
    prm%kernel = DUST_KERNEL_ROSENHEAD
    prm%use_divfilt = merge(1_c_int, 0_c_int, gpu_sim_param%use_divfilt)
    prm%reserved = 0
    octree%nlevels = nlevels
    octree%lvl_solid = gpu_sim_param%lvl_solid
    if (octree%lvl_solid < 1) octree%lvl_solid = max(nlevels - 2, 1)
    ! the device capacity grows with the first sort that needs more
    octree%prm = prm
    octree%cap = capacity
    call check(dust_fmm_create(prm, int(capacity, c_int), octree%gpu), 'initialize_octree')
    allocate(octree%slot(capacity), octree%cen(3, capacity), octree%dir(3, capacity), &
             octree%mag(capacity), octree%rv(capacity))
    octree%n = 0
  end subroutine initialize_octree

  subroutine destroy_octree(octree)
    type(t_octree), intent(inout) :: octree
    if (c_associated(octree%gpu)) call check(dust_fmm_destroy(octree%gpu), 'destroy_octree')
    octree%gpu = c_null_ptr
    if (allocated(octree%slot)) deallocate(octree%slot, octree%cen, octree%dir, octree%mag, &
                                           octree%rv)
  end subroutine destroy_octree

  ! A handle whose capacity holds n particles (same parameters; the capacity
  ! at least doubles, so that a growing wake rebuilds it rarely).
  subroutine ensure_capacity(octree, n)
    type(t_octree), intent(inout) :: octree
    integer, intent(in) :: n
    if (n <= octree%cap) return
    octree%cap = max(n, 2 * octree%cap)
    deallocate(octree%slot, octree%cen, octree%dir, octree%mag, octree%rv)
    call check(dust_fmm_destroy(octree%gpu), 'sort_particles')
    call check(dust_fmm_create(octree%prm, int(octree%cap, c_int), octree%gpu), 'sort_particles')
    allocate(octree%slot(octree%cap), octree%cen(3, octree%cap), octree%dir(3, octree%cap), &
             octree%mag(octree%cap), octree%rv(octree%cap))
  end subroutine ensure_capacity

  ! sort_particles: the first n_prt non-free slots of wparts in slot order
  ! (the reference's visiting loop, mod_octree.f90:746-773); with
  ! particles_redistribution the freed particles get free = .true., the kept
  ! ones their new strengths, and n_prt is lowered (827-894).
  subroutine sort_particles(wparts, n_prt, elem, octree)
    type(t_vortparts), intent(inout) :: wparts
    integer, intent(inout) :: n_prt
    type(t_elems), intent(in) :: elem
    type(t_octree), intent(inout) :: octree
    integer :: ip, ipp, n_kept
    integer(c_int) :: nl, nk
    integer(c_int8_t), allocatable :: keep(:)
    real(wp) :: nocen(3, 1)
    call ensure_capacity(octree, n_prt)
    ipp = 0
    do ip = 1, n_prt
      ipp = ipp + 1
      do
        if (ipp > size(wparts%free)) then
          print '(a)', 'mod_octree_gpu: not enough non-free particles while sorting'
          error stop 1
        end if
        if (.not. wparts%free(ipp)) exit
        ipp = ipp + 1
      end do
      octree%slot(ip) = ipp
      octree%cen(:, ip) = wparts%cen(:, ipp)
      octree%dir(:, ip) = wparts%dir(:, ipp)
      octree%mag(ip) = wparts%mag(ipp)
      octree%rv(ip) = wparts%r_Vortex(ipp)
    end do
    octree%n = n_prt
    call check(dust_fmm_set_particles(octree%gpu, int(n_prt, c_int), octree%cen, octree%dir, &
                                      octree%mag, octree%rv), 'sort_particles')
    if (gpu_sim_param%use_pr) then
      if (elem%n > 0) then
        call check(dust_fmm_set_redistribution(octree%gpu, 1_c_int, int(octree%lvl_solid, c_int), &
                   gpu_sim_param%part_redist_ratio, int(elem%n, c_int), elem%cen), &
                   'sort_particles')
      else
        call check(dust_fmm_set_redistribution(octree%gpu, 1_c_int, int(octree%lvl_solid, c_int), &
                   gpu_sim_param%part_redist_ratio, 0_c_int, nocen), 'sort_particles')
      end if
    else
      call check(dust_fmm_set_redistribution(octree%gpu, 0_c_int, 1_c_int, 1.0_wp, 0_c_int, &
                 nocen), 'sort_particles')
    end if
    call check(dust_fmm_sort_particles(octree%gpu, nl), 'sort_particles')
    octree%nleaves = nl
    if (gpu_sim_param%use_pr .and. n_prt > 0) then
      allocate(keep(n_prt))
      call check(dust_fmm_redistributed(octree%gpu, nk, keep, octree%dir, octree%mag), &
                 'sort_particles')
      n_kept = nk
      do ip = 1, n_prt
        ipp = octree%slot(ip)
        wparts%free(ipp) = (keep(ip) == 0_c_int8_t)
        wparts%dir(:, ipp) = octree%dir(:, ip)
        wparts%mag(ipp) = octree%mag(ip)
      end do
      n_prt = n_kept
    end if
  end subroutine sort_particles

  ! calculate_multipole: part is unused, as in the reference
  subroutine calculate_multipole(part, octree)
    type(t_vortparts), intent(in) :: part
    type(t_octree), intent(inout) :: octree
    call check(dust_fmm_calculate_multipole(octree%gpu), 'calculate_multipole')
  end subroutine calculate_multipole

  ! apply_multipole: particle terms and wind on the device, then the element
  ! loops of elem, wpan, wrin (no stretching) and wvort, written into the
  ! slots of the particles the last sort kept
  subroutine apply_multipole(part, octree, elem, wpan, wrin, wvort)
    type(t_vortparts), intent(inout) :: part
    type(t_octree), intent(inout) :: octree
    type(t_elems), intent(in), target :: elem, wpan, wrin, wvort
    real(wp), allocatable, target :: vel(:,:), stretch(:,:), tv(:), alpha(:,:), ru(:,:)
    type(t_dust_elem_group) :: grp(4)
    type(t_dust_elem_params) :: eprm
    integer :: ip, ipp, n, ng
    logical :: vs
    n = octree%n
    allocate(vel(3, max(n, 1)), stretch(3, max(n, 1)), tv(max(n, 1)), alpha(3, max(n, 1)))
    call check(dust_fmm_apply_multipole(octree%gpu, vel, stretch, tv), 'apply_multipole')
    if (octree%prm%use_divfilt /= 0 .and. octree%prm%use_vs /= 0) then
      allocate(ru(3, max(n, 1)))
      call check(dust_fmm_rotu(octree%gpu, ru), 'apply_multipole')
    end if
    ng = 0
    call add_group(elem, 1, gpu_sim_param%vs_elems)
    call add_group(wpan, 1, gpu_sim_param%vs_elems)
    call add_group(wrin, 0, .false.)
    call add_group(wvort, 1, gpu_sim_param%vs_elems)
    vs = gpu_sim_param%use_vs .and. gpu_sim_param%vs_elems
    if (ng > 0 .and. n > 0) then
      eprm%ff_ratio_doublet = gpu_sim_param%FarFieldRatioDoublet
      eprm%rankine_rad = gpu_sim_param%RankineRad
      eprm%cutoff_rad = gpu_sim_param%CutoffRad
      eprm%vortex_rad = gpu_sim_param%VortexRad
      eprm%precision = merge(DUST_ELEM_TUNED, DUST_ELEM_FP64, gpu_sim_param%tuned)
      eprm%reserved = 0
      do ip = 1, n
        alpha(:, ip) = octree%mag(ip) * octree%dir(:, ip)
      end do
      if (vs) then
        call check(dust_elem_apply(int(ng, c_int), grp, eprm, int(n, c_int), octree%cen, &
                   c_loc(alpha), vel, c_loc(stretch)), 'apply_multipole')
      else
        call check(dust_elem_apply(int(ng, c_int), grp, eprm, int(n, c_int), octree%cen, &
                   c_null_ptr, vel, c_null_ptr), 'apply_multipole')
      end if
    end if
    do ip = 1, n
      ipp = octree%slot(ip)
      if (part%free(ipp)) cycle
      part%vel(:, ipp) = vel(:, ip)
      part%stretch(:, ipp) = stretch(:, ip)
      part%turbvisc(ipp) = tv(ip)
      if (allocated(ru) .and. allocated(part%rotu)) part%rotu(:, ipp) = ru(:, ip)
    end do
  contains
    subroutine add_group(e, st, want)
      type(t_elems), intent(in), target :: e
      integer, intent(in) :: st
      logical, intent(in) :: want
      if (e%n <= 0) return
      ng = ng + 1
      grp(ng)%kind = int(e%kind, c_int)
      grp(ng)%n = int(e%n, c_int)
      grp(ng)%n_ver = c_null_ptr
      if (allocated(e%n_ver)) grp(ng)%n_ver = c_loc(e%n_ver)
      grp(ng)%ver = c_loc(e%ver)
      grp(ng)%mag = c_loc(e%mag)
      grp(ng)%sigma = c_null_ptr
      if (allocated(e%sigma)) grp(ng)%sigma = c_loc(e%sigma)
      grp(ng)%stretch = merge(int(st, c_int), 0_c_int, want)
      grp(ng)%reserved = 0
    end subroutine add_group
  end subroutine apply_multipole

  ! apply_multipole_panels: elem%uvort += particle velocity at elem%cen
  subroutine apply_multipole_panels(octree, elem)
    type(t_octree), intent(inout) :: octree
    type(t_elems), intent(inout) :: elem
    if (elem%n <= 0) return
    call check(dust_fmm_apply_multipole_panels(octree%gpu, int(elem%n, c_int), elem%cen, &
                                               elem%uvort), 'apply_multipole_panels')
  end subroutine apply_multipole_panels

end module mod_octree_gpu
