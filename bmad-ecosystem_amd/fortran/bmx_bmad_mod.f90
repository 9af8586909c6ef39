!+
! Module bmx_bmad_mod
!
! iso_c_binding layer between a Fortran tracking host and the MI355X tracking library
! (include/bmx.h). It keeps the calling convention of Bmad's beam tracking:
!
!   track_beam / track_bunch  (bmad/multiparticle/beam_mod.f90:40,103): a bunch of particles,
!       each a coord_struct (bmad/modules/bmad_struct.f90:551-582), tracked through the
!       lattice branch; particle loss is recorded in %state, never reported as an error
!       (bmad/code/track1.f90:39-41).
!   track_all (bmad/code/track_all.f90:25): one particle through the elements, orbit(0:n)
!       at every element end, track_state = index of the element where it was lost.
!
! The library works on structure-of-arrays FP64 buffers. This module converts the host's
! particle records to SoA and back field by field: coord_struct%t is real(qp), so the record
! layout depends on the compiler and is never copied as raw memory.
!
! bmx_coord_struct carries the coord_struct components the path reads and writes, with the
! same names and meanings; a host built on Bmad assigns them from/to its coord_struct array
! (the adapter is a loop of component assignments).
!-

module bmx_bmad_mod

use, intrinsic :: iso_c_binding
implicit none

private

integer, parameter, public :: bmx_rp = c_double

! Status codes (bmx.h)
integer(c_int32_t), parameter, public :: bmx_ok = 0, bmx_err_arg = 1, bmx_err_unsupported = 2, &
    bmx_err_hip = 3, bmx_err_p0c = 4, bmx_err_nomem = 5, bmx_err_budget = 6, bmx_err_state = 7, &
    bmx_err_not_found = 8

! coord_struct%state codes (bmad_struct.f90:527-532)
integer, parameter, public :: bmx_not_set = -999, bmx_alive = 1, bmx_lost = 2, &
    bmx_lost_neg_x = 3, bmx_lost_pos_x = 4, bmx_lost_neg_y = 5, bmx_lost_pos_y = 6, &
    bmx_lost_z = 7, bmx_lost_pz = 8

! track_all's track_state when no particle was lost (bmad_struct.f90: moving_forward$)
integer, parameter, public :: bmx_moving_forward = -1

! Element keys (bmad_struct.f90:1696-1710)
integer(c_int32_t), parameter, public :: bmx_drift = 1, bmx_sbend = 2, bmx_quadrupole = 3, &
    bmx_sextupole = 5, bmx_rfcavity = 9, bmx_marker = 14, bmx_kicker = 15, bmx_octupole = 17, &
    bmx_multipole = 19, bmx_ab_multipole = 22, &
    bmx_monitor = 31, bmx_instrument = 32, bmx_hkicker = 33, bmx_vkicker = 34, &
    bmx_rcollimator = 35, bmx_ecollimator = 36, bmx_pipe = 44, bmx_thick_multipole = 67

integer, parameter, public :: bmx_n_pole = 22

! bmx_ele of bmx.h: the raw ele%value attributes of one element (bmad_struct.f90:1469).
type, bind(c), public :: bmx_ele_struct
  integer(c_int32_t) :: key = 0, is_on = 1, fringe_type = 0, fringe_at = 0
  integer(c_int32_t) :: aperture_type = 0, aperture_at = 0, offset_moves_aperture = 0, n_multipole = 0
  real(c_double) :: l = 0
  real(c_double) :: k1 = 0, k2 = 0, k3 = 0
  real(c_double) :: g = 0, dg = 0, e1 = 0, e2 = 0, fint = 0, fintx = 0, hgap = 0, hgapx = 0
  real(c_double) :: hkick = 0, vkick = 0, kick = 0
  real(c_double) :: tilt = 0, ref_tilt = 0
  real(c_double) :: x_offset = 0, y_offset = 0, z_offset = 0, x_pitch = 0, y_pitch = 0
  real(c_double) :: voltage = 0, rf_frequency = 0, phi0 = 0, phi0_multipass = 0
  real(c_double) :: x1_limit = 0, x2_limit = 0, y1_limit = 0, y2_limit = 0
  real(c_double) :: px_aperture_width2 = 0, px_aperture_center = 0
  real(c_double) :: py_aperture_width2 = 0, py_aperture_center = 0
  real(c_double) :: z_aperture_width2 = 0, z_aperture_center = 0
  real(c_double) :: pz_aperture_width2 = 0, pz_aperture_center = 0
  real(c_double) :: ds_step = 0
  real(c_double) :: p0c = 0, e_tot = 0
  real(c_double) :: r0_mag = 0
  real(c_double) :: a_pole(0:bmx_n_pole-1) = 0
  real(c_double) :: b_pole(0:bmx_n_pole-1) = 0
  real(c_double) :: r0_elec = 0
  real(c_double) :: a_pole_elec(0:bmx_n_pole-1) = 0
  real(c_double) :: b_pole_elec(0:bmx_n_pole-1) = 0
  integer(c_int32_t) :: n_multipole_elec = 0, exact_multipoles = 0
  real(c_double) :: fq1 = 0, fq2 = 0
end type

! bmx_beam_view of bmx.h
type, bind(c) :: bmx_beam_view
  type(c_ptr) :: x = c_null_ptr, px = c_null_ptr, y = c_null_ptr, py = c_null_ptr
  type(c_ptr) :: z = c_null_ptr, pz = c_null_ptr, beta = c_null_ptr
  type(c_ptr) :: state = c_null_ptr, loss_ele = c_null_ptr, loss_turn = c_null_ptr
end type

! bmx_rad_map (include/bmx.h): one half of an element's rad_map_struct (bmad_struct.f90:988-994)
! without xfer_damp_mat. The C matrices are row-major: fill damp_dmat and stoc_mat with the
! transposes of the rad_map_struct components.
type, bind(c), public :: bmx_rad_map_c
  real(c_double) :: ref_orb(6) = 0
  real(c_double) :: damp_dmat(6,6) = 0
  real(c_double) :: xfer_damp_vec(6) = 0
  real(c_double) :: stoc_mat(6,6) = 0
end type

! bmx_rad_config: the bmad_com radiation switches, synch_rad_scale and the fluctuation seed.
type, bind(c), public :: bmx_rad_config_c
  integer(c_int32_t) :: damping_on = 0
  integer(c_int32_t) :: fluctuations_on = 0
  integer(c_int32_t) :: zero_average = 0
  integer(c_int32_t) :: reserved = 0
  real(c_double) :: synch_rad_scale = 1
  integer(c_int64_t) :: seed = 0
end type

! The coord_struct components on the path (bmad_struct.f90:551-582).
type, public :: bmx_coord_struct
  real(bmx_rp) :: vec(6) = 0
  real(bmx_rp) :: p0c = 0
  real(bmx_rp) :: beta = -1
  integer :: ix_ele = -1
  integer :: ix_turn = 0
  integer :: state = bmx_not_set
end type

! One device context plus one compiled lattice branch.
type, public :: bmx_lat_struct
  type(c_ptr) :: ctx = c_null_ptr
  type(c_ptr) :: lat = c_null_ptr
  integer :: n_ele = 0
  real(bmx_rp) :: p0c = 0
end type

interface
  function c_ctx_create(device, ctx) bind(c, name = 'bmx_ctx_create') result(st)
    import :: c_int32_t, c_ptr
    integer(c_int32_t), value :: device
    type(c_ptr) :: ctx
    integer(c_int32_t) :: st
  end function

  function c_ctx_destroy(ctx) bind(c, name = 'bmx_ctx_destroy') result(st)
    import :: c_int32_t, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int32_t) :: st
  end function

  function c_ctx_synchronize(ctx) bind(c, name = 'bmx_ctx_synchronize') result(st)
    import :: c_int32_t, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int32_t) :: st
  end function

  ! bmx_lattice_create with radiation kicks: maps(1, i), maps(2, i) = rm0, rm1 of element i.
  function bmx_lattice_create_rad(ctx, eles, n_ele, mc2, charge, maps, config, lat) &
      bind(c, name = 'bmx_lattice_create_rad') result(st)
    import :: c_int32_t, c_ptr, c_double, bmx_rad_config_c
    type(c_ptr), value :: ctx, eles
    integer(c_int32_t), value :: n_ele
    real(c_double), value :: mc2
    integer(c_int32_t), value :: charge
    type(c_ptr), value :: maps
    type(bmx_rad_config_c), intent(in) :: config
    type(c_ptr) :: lat
    integer(c_int32_t) :: st
  end function

  function bmx_lattice_set_radiation(lat, config) bind(c, name = 'bmx_lattice_set_radiation') result(st)
    import :: c_int32_t, c_ptr, bmx_rad_config_c
    type(c_ptr), value :: lat
    type(bmx_rad_config_c), intent(in) :: config
    integer(c_int32_t) :: st
  end function

  function c_lattice_create(ctx, eles, n_ele, mc2, charge, lat) bind(c, name = 'bmx_lattice_create') result(st)
    import :: c_int32_t, c_ptr, c_double, bmx_ele_struct
    type(c_ptr), value :: ctx
    type(bmx_ele_struct), intent(in) :: eles(*)
    integer(c_int32_t), value :: n_ele
    real(c_double), value :: mc2
    integer(c_int32_t), value :: charge
    type(c_ptr) :: lat
    integer(c_int32_t) :: st
  end function

  function c_closed_orbit(ctx, lats, n_lat, n_dim, orbit, mat, max_iter, rel_tol, abs_tol, status) &
                          bind(c, name = 'bmx_closed_orbit') result(st)
    import :: c_int32_t, c_ptr, c_double
    type(c_ptr), value :: ctx
    type(c_ptr), intent(in) :: lats(*)
    integer(c_int32_t), value :: n_lat, n_dim, max_iter
    real(c_double), intent(inout) :: orbit(*)
    real(c_double), intent(out) :: mat(*)
    real(c_double), value :: rel_tol, abs_tol
    integer(c_int32_t), intent(out) :: status(*)
    integer(c_int32_t) :: st
  end function

  function c_beam_file_write(path, host, n, species, charge, p0c, ix_ele, append) &
                             bind(c, name = 'bmx_beam_file_write') result(st)
    import :: c_int32_t, c_int64_t, c_char, c_double, bmx_beam_view
    character(kind = c_char), intent(in) :: path(*), species(*)
    type(bmx_beam_view), intent(in) :: host
    integer(c_int64_t), value :: n
    real(c_double), value :: charge, p0c
    integer(c_int32_t), value :: ix_ele, append
    integer(c_int32_t) :: st
  end function

  function c_lattice_specialize(lat) bind(c, name = 'bmx_lattice_specialize') result(st)
    import :: c_int32_t, c_ptr
    type(c_ptr), value :: lat
    integer(c_int32_t) :: st
  end function

  function c_lattice_destroy(lat) bind(c, name = 'bmx_lattice_destroy') result(st)
    import :: c_int32_t, c_ptr
    type(c_ptr), value :: lat
    integer(c_int32_t) :: st
  end function

  function c_beam_create(ctx, n, beam) bind(c, name = 'bmx_beam_create') result(st)
    import :: c_int32_t, c_int64_t, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int64_t), value :: n
    type(c_ptr) :: beam
    integer(c_int32_t) :: st
  end function

  function c_beam_upload(beam, host, n) bind(c, name = 'bmx_beam_upload') result(st)
    import :: c_int32_t, c_int64_t, c_ptr, bmx_beam_view
    type(c_ptr), value :: beam
    type(bmx_beam_view), intent(in) :: host
    integer(c_int64_t), value :: n
    integer(c_int32_t) :: st
  end function

  function c_beam_download(beam, host, n) bind(c, name = 'bmx_beam_download') result(st)
    import :: c_int32_t, c_int64_t, c_ptr, bmx_beam_view
    type(c_ptr), value :: beam
    type(bmx_beam_view), intent(in) :: host
    integer(c_int64_t), value :: n
    integer(c_int32_t) :: st
  end function

  function c_beam_destroy(beam) bind(c, name = 'bmx_beam_destroy') result(st)
    import :: c_int32_t, c_ptr
    type(c_ptr), value :: beam
    integer(c_int32_t) :: st
  end function

  function c_track_turns(ctx, lat, beam, n_turns, turn0) bind(c, name = 'bmx_track_turns') result(st)
    import :: c_int32_t, c_ptr
    type(c_ptr), value :: ctx, lat, beam
    integer(c_int32_t), value :: n_turns, turn0
    integer(c_int32_t) :: st
  end function

  function c_track_element(ctx, lat, beam, ix_ele, turn) bind(c, name = 'bmx_track_element') result(st)
    import :: c_ptr, c_int32_t
    type(c_ptr), value :: ctx, lat, beam
    integer(c_int32_t), value :: ix_ele, turn
    integer(c_int32_t) :: st
  end function

  function c_track_bunch(ctx, lat, beam, ix_ele1, ix_ele2, turn) bind(c, name = 'bmx_track_bunch') result(st)
    import :: c_ptr, c_int32_t
    type(c_ptr), value :: ctx, lat, beam
    integer(c_int32_t), value :: ix_ele1, ix_ele2, turn
    integer(c_int32_t) :: st
  end function

  function c_track_orbit(ctx, lat, beam, ix_start, ix_end, turn, orbit) bind(c, name = 'bmx_track_orbit') result(st)
    import :: c_int32_t, c_ptr
    type(c_ptr), value :: ctx, lat, beam
    integer(c_int32_t), value :: ix_start, ix_end, turn
    type(c_ptr), value :: orbit
    integer(c_int32_t) :: st
  end function

  function c_last_error() bind(c, name = 'bmx_last_error') result(msg)
    import :: c_ptr
    type(c_ptr) :: msg
  end function

  function c_strlen(s) bind(c, name = 'strlen') result(n)
    import :: c_ptr, c_size_t
    type(c_ptr), value :: s
    integer(c_size_t) :: n
  end function
end interface

! A bunch as the hook and span entry points take it: the particles, and their copy on the
! device, which stays allocated between calls (bmx_bunch_free releases it).
type, public :: bmx_bunch_struct
  type(bmx_coord_struct), allocatable :: particle(:)
  type(c_ptr) :: beam = c_null_ptr
  integer(c_int64_t) :: n_device = 0
end type

public :: bmx_lat_init, bmx_lat_free, bmx_track_bunch, bmx_track_all, bmx_last_error_message
public :: bmx_closed_orbit_calc, bmx_write_beam_file
public :: bmx_track1_bunch_hook, bmx_track_bunch_span, bmx_bunch_free
public :: bmx_lattice_create_rad, bmx_lattice_set_radiation

contains

!+
! Function bmx_last_error_message() result(msg)
! Message of the last failed library call on this thread (bmx_last_error).
!-
function bmx_last_error_message() result(msg)
  character(:), allocatable :: msg
  type(c_ptr) :: p
  character(kind=c_char), pointer :: chars(:)
  integer :: i, n
  p = c_last_error()
  if (.not. c_associated(p)) then
    msg = ''
    return
  endif
  n = int(c_strlen(p))
  call c_f_pointer(p, chars, [n])
  allocate(character(n) :: msg)
  do i = 1, n
    msg(i:i) = chars(i)
  enddo
end function

!+
! Subroutine bmx_lat_init (lat, eles, mc2, charge, err_flag, device, specialize)
!
! Compiles the element table eles(1:n) (the tracking part of a lattice branch, in order) on
! device `device` (default 0). mc2 [eV] and charge are the species' (particle_species_mod.f90).
! err_flag is set, and the message is in bmx_last_error_message(), if an element is outside
! the supported set or the reference momenta differ. specialize (default false): also compile
! tracking kernels for the lattice's repeated cell (bmx_lattice_specialize); a lattice without
! one keeps the generic kernel.
!-
subroutine bmx_lat_init (lat, eles, mc2, charge, err_flag, device, specialize)
  type(bmx_lat_struct), intent(inout) :: lat
  type(bmx_ele_struct), intent(in) :: eles(:)
  real(bmx_rp), intent(in) :: mc2
  integer, intent(in) :: charge
  logical, intent(out) :: err_flag
  integer, optional, intent(in) :: device
  logical, optional, intent(in) :: specialize
  integer(c_int32_t) :: dev, st

  err_flag = .true.
  dev = 0
  if (present(device)) dev = device
  if (.not. c_associated(lat%ctx)) then
    st = c_ctx_create(dev, lat%ctx)
    if (st /= bmx_ok) return
  endif
  if (c_associated(lat%lat)) then
    st = c_lattice_destroy(lat%lat)
    lat%lat = c_null_ptr
  endif
  st = c_lattice_create(lat%ctx, eles, int(size(eles), c_int32_t), real(mc2, c_double), &
                        int(charge, c_int32_t), lat%lat)
  if (st /= bmx_ok) return
  if (present(specialize)) then
    if (specialize) then
      st = c_lattice_specialize(lat%lat)
      if (st /= bmx_ok .and. st /= bmx_err_unsupported) return
    endif
  endif
  lat%n_ele = size(eles)
  lat%p0c = 0
  if (size(eles) > 0) lat%p0c = eles(1)%p0c
  err_flag = .false.
end subroutine

subroutine bmx_lat_free (lat)
  type(bmx_lat_struct), intent(inout) :: lat
  integer(c_int32_t) :: st
  if (c_associated(lat%lat)) st = c_lattice_destroy(lat%lat)
  if (c_associated(lat%ctx)) st = c_ctx_destroy(lat%ctx)
  lat%lat = c_null_ptr
  lat%ctx = c_null_ptr
  lat%n_ele = 0
end subroutine

!+
! Subroutine bmx_track_bunch (lat, particle, n_turns, err_flag, ix_turn0)
!
! track_bunch (beam_mod.f90:103) from the start of the branch to its end, n_turns times
! (the turn loop of ltt_run_beam_mode, lt_tracking_mod.f90:1631-1716), for particle(:).
! Particles not alive on entry are left untouched. A particle lost on the way keeps the
! coordinates it had when lost, gets Bmad's loss code in %state, the 1-based index of the
! element it was lost in in %ix_ele and the 0-based turn in %ix_turn. Survivors end with
! %ix_ele = n_ele and %ix_turn = ix_turn0 + n_turns - 1.
!-
subroutine bmx_track_bunch (lat, particle, n_turns, err_flag, ix_turn0)
  type(bmx_lat_struct), intent(in) :: lat
  type(bmx_coord_struct), intent(inout) :: particle(:)
  integer, intent(in) :: n_turns
  logical, intent(out) :: err_flag
  integer, optional, intent(in) :: ix_turn0

  real(c_double), allocatable, target :: x(:), px(:), y(:), py(:), z(:), pz(:), beta(:)
  integer(c_int32_t), allocatable, target :: state(:), loss_ele(:), loss_turn(:)
  type(bmx_beam_view) :: view
  type(c_ptr) :: beam
  integer(c_int32_t) :: st, st2, t0
  integer(c_int64_t) :: n
  integer :: i

  err_flag = .true.
  n = size(particle, kind = c_int64_t)
  t0 = 0
  if (present(ix_turn0)) t0 = ix_turn0
  allocate(x(n), px(n), y(n), py(n), z(n), pz(n), beta(n), state(n), loss_ele(n), loss_turn(n))
  do i = 1, int(n)
    x(i) = particle(i)%vec(1); px(i) = particle(i)%vec(2)
    y(i) = particle(i)%vec(3); py(i) = particle(i)%vec(4)
    z(i) = particle(i)%vec(5); pz(i) = particle(i)%vec(6)
    beta(i) = particle(i)%beta
    state(i) = particle(i)%state
    loss_ele(i) = 0
    loss_turn(i) = -1
  enddo
  call fill_view(view, x, px, y, py, z, pz, beta, state, loss_ele, loss_turn)

  beam = c_null_ptr
  st = c_beam_create(lat%ctx, n, beam)
  if (st /= bmx_ok) return
  st = c_beam_upload(beam, view, n)
  if (st == bmx_ok) st = c_track_turns(lat%ctx, lat%lat, beam, int(n_turns, c_int32_t), t0)
  if (st == bmx_ok) st = c_beam_download(beam, view, n)
  st2 = c_beam_destroy(beam)
  if (st /= bmx_ok .or. st2 /= bmx_ok) return

  do i = 1, int(n)
    if (particle(i)%state /= bmx_alive) cycle
    particle(i)%vec = [x(i), px(i), y(i), py(i), z(i), pz(i)]
    particle(i)%beta = beta(i)
    particle(i)%state = state(i)
    if (state(i) == bmx_alive) then
      particle(i)%ix_ele = lat%n_ele
      particle(i)%ix_turn = t0 + n_turns - 1
    else
      particle(i)%ix_ele = loss_ele(i)
      particle(i)%ix_turn = loss_turn(i)
    endif
  enddo
  err_flag = .false.
end subroutine

!+
! Subroutine bmx_track_all (lat, orbit, track_state, err_flag)
!
! track_all (track_all.f90:25-145) for one particle: orbit(0) is the start, on return
! orbit(i) holds the particle at the exit of element i for every element it passed. If the
! particle is lost in element i: track_state = i, orbit(i) holds the loss state (with zero
! coordinates if it was lost at the element's entrance), orbit(i+1:) are reset to
! bmx_coord_struct() and err_flag is set, as track_all does (:81-88). Otherwise
! track_state = bmx_moving_forward. A library failure sets err_flag with track_state =
! bmx_moving_forward and the message in bmx_last_error_message().
!-
subroutine bmx_track_all (lat, orbit, track_state, err_flag)
  type(bmx_lat_struct), intent(in) :: lat
  type(bmx_coord_struct), intent(inout) :: orbit(0:)
  integer, intent(out) :: track_state
  logical, intent(out) :: err_flag

  real(c_double), target :: x(1), px(1), y(1), py(1), z(1), pz(1), beta(1)
  integer(c_int32_t), target :: state(1), loss_ele(1), loss_turn(1)
  real(c_double), allocatable, target :: rec(:, :)
  type(bmx_beam_view) :: view
  type(c_ptr) :: beam
  integer(c_int32_t) :: st, st2
  integer :: i, n_ele

  err_flag = .true.
  track_state = bmx_moving_forward
  n_ele = lat%n_ele
  if (ubound(orbit, 1) < n_ele) return
  x = orbit(0)%vec(1); px = orbit(0)%vec(2); y = orbit(0)%vec(3)
  py = orbit(0)%vec(4); z = orbit(0)%vec(5); pz = orbit(0)%vec(6)
  beta = orbit(0)%beta
  state = orbit(0)%state
  loss_ele = 0
  loss_turn = -1
  call fill_view(view, x, px, y, py, z, pz, beta, state, loss_ele, loss_turn)
  ! rec(1:8, i): x, px, y, py, z, pz, beta, state at the exit of element i (i = 1..n_ele)
  allocate(rec(8, n_ele))
  rec = 0

  beam = c_null_ptr
  st = c_beam_create(lat%ctx, 1_c_int64_t, beam)
  if (st /= bmx_ok) return
  st = c_beam_upload(beam, view, 1_c_int64_t)
  if (st == bmx_ok) st = c_track_orbit(lat%ctx, lat%lat, beam, 1_c_int32_t, int(n_ele, c_int32_t), &
                                         int(orbit(0)%ix_turn, c_int32_t), c_loc(rec))
  if (st == bmx_ok) st = c_beam_download(beam, view, 1_c_int64_t)
  st2 = c_beam_destroy(beam)
  if (st /= bmx_ok .or. st2 /= bmx_ok) return

  err_flag = .false.
  do i = 1, n_ele
    if (track_state /= bmx_moving_forward) then
      orbit(i) = bmx_coord_struct()
      cycle
    endif
    orbit(i)%vec = rec(1:6, i)
    orbit(i)%beta = rec(7, i)
    orbit(i)%state = nint(rec(8, i))
    orbit(i)%p0c = orbit(0)%p0c
    orbit(i)%ix_ele = i
    orbit(i)%ix_turn = orbit(0)%ix_turn
    if (orbit(i)%state /= bmx_alive) then
      track_state = i
      err_flag = .true.
    endif
  enddo
end subroutine

!+
! Subroutine bmx_closed_orbit_calc (lat, closed_orb, i_dim, err_flag, mat6, rel_tol, abs_tol)
!
! closed_orbit_calc (bmad/code/closed_orbit_calc.f90:192) on the GPU: closed_orb%vec is the
! starting guess on entry (its pz is the energy for i_dim = 4, which closes x, px, y, py with
! the rf off) and the closed orbit on return (i_dim = 6 closes all six with the rf on). mat6
! (optional) receives the one-turn matrix at the orbit. The search ends when every Newton step
! is below |orbit| rel_tol + abs_tol (defaults 1e-8 and 1e-11, bmad_com's rel_tol_tracking and
! abs_tol_tracking). err_flag is set if it does not converge in 100 steps or a probe particle
! is lost.
!-
subroutine bmx_closed_orbit_calc (lat, closed_orb, i_dim, err_flag, mat6, rel_tol, abs_tol)
  type(bmx_lat_struct), intent(in) :: lat
  type(bmx_coord_struct), intent(inout) :: closed_orb
  integer, intent(in) :: i_dim
  logical, intent(out) :: err_flag
  real(bmx_rp), optional, intent(out) :: mat6(6, 6)
  real(bmx_rp), optional, intent(in) :: rel_tol, abs_tol

  type(c_ptr) :: lats(1)
  real(c_double) :: orb(6), m(6, 6), rt, at
  integer(c_int32_t) :: st, status(1)

  err_flag = .true.
  rt = 1e-8_c_double
  at = 1e-11_c_double
  if (present(rel_tol)) rt = rel_tol
  if (present(abs_tol)) at = abs_tol
  lats(1) = lat%lat
  orb = closed_orb%vec
  st = c_closed_orbit(lat%ctx, lats, 1_c_int32_t, int(i_dim, c_int32_t), orb, m, 100_c_int32_t, &
                      rt, at, status)
  if (st /= bmx_ok) return
  closed_orb%vec = orb
  ! the library's matrix is row-major: m(c, r) = d out_r / d in_c
  if (present(mat6)) mat6 = transpose(m)
  err_flag = status(1) /= 0
end subroutine

!+
! Subroutine bmx_write_beam_file (file_name, particle, species, charge, ix_ele, err_flag)
!
! write_beam_file (bmad/multiparticle/beam_file_io.f90:29) for one bunch in the column layout:
! vec, p0c, beta, state and ix_ele of every particle (ix_ele for live ones, the particle's ix_ele
! -- its loss element -- for lost ones), charge per particle and species as parameters; 17
! significant digits, so the file is a bit-identical checkpoint.
!-
subroutine bmx_write_beam_file (file_name, particle, species, charge, ix_ele, err_flag)
  character(*), intent(in) :: file_name, species
  type(bmx_coord_struct), intent(in) :: particle(:)
  real(bmx_rp), intent(in) :: charge
  integer, intent(in) :: ix_ele
  logical, intent(out) :: err_flag

  real(c_double), allocatable, target :: x(:), px(:), y(:), py(:), z(:), pz(:), beta(:)
  integer(c_int32_t), allocatable, target :: state(:), loss_ele(:), loss_turn(:)
  type(bmx_beam_view) :: view
  integer(c_int32_t) :: st
  integer :: n

  n = size(particle)
  allocate(x(n), px(n), y(n), py(n), z(n), pz(n), beta(n), state(n), loss_ele(n), loss_turn(n))
  x = particle%vec(1); px = particle%vec(2); y = particle%vec(3)
  py = particle%vec(4); z = particle%vec(5); pz = particle%vec(6)
  beta = particle%beta
  state = particle%state
  loss_ele = particle%ix_ele
  loss_turn = particle%ix_turn
  call fill_view(view, x, px, y, py, z, pz, beta, state, loss_ele, loss_turn)
  st = c_beam_file_write(trim(file_name) // c_null_char, view, int(n, c_int64_t), &
                         trim(species) // c_null_char, real(charge, c_double), &
                         real(particle(1)%p0c, c_double), int(ix_ele, c_int32_t), 0_c_int32_t)
  err_flag = st /= bmx_ok
end subroutine

!+
! Subroutine bmx_track1_bunch_hook (lat, bunch, ix_ele, err, finished, turn, direction)
!
! The body of a track1_bunch_hook (bmad_routine_interface.f90:3937, track1_bunch_hook_def;
! called once per element per bunch at beam_mod.f90:263-266): tracks bunch%particle through
! element ix_ele (1-based, ele%ix_ele of the hook's ele) with bmx_track_element (the GPU form of
! track1_bunch_hom, beam_utils.f90:32-90) and sets finished = .true., so Bmad skips its own
! tracking of that element. A Bmad hook installed in track1_bunch_hook_ptr calls it with its
! ele%ix_ele and the bunch's particles copied component by component. Particles not alive on
! entry are skipped; a particle lost in the element gets the loss code in %state, ix_ele and
! turn in %ix_ele and %ix_turn; the others get %ix_ele = ix_ele. direction must be +1 (the
! default). err (and finished = .false.) on a library failure, with the message in
! bmx_last_error_message().
!-
subroutine bmx_track1_bunch_hook (lat, bunch, ix_ele, err, finished, turn, direction)
  type(bmx_lat_struct), intent(in) :: lat
  type(bmx_bunch_struct), intent(inout) :: bunch
  integer, intent(in) :: ix_ele
  logical, intent(out) :: err, finished
  integer, optional, intent(in) :: turn, direction
  integer(c_int32_t) :: t

  err = .true.
  finished = .false.
  if (present(direction)) then
    if (direction /= 1) return
  endif
  t = 0
  if (present(turn)) t = turn
  call bunch_through_device(lat, bunch, err, 1, int(ix_ele, c_int32_t), 0_c_int32_t, t)
  finished = .not. err
end subroutine

!+
! Subroutine bmx_track_bunch_span (lat, bunch, ele1, ele2, err, turn)
!
! track_bunch (beam_mod.f90:103) with its argument list (lat, bunch, ele1, ele2, err): the
! bunch at the end of element ele1 (0..n_ele; ele1 itself is not tracked) to the end of element
! ele2, i.e. elements ele1+1 .. ele2, or ele1+1 .. n_ele then 1 .. ele2 when ele2 <= ele1
! (beam_mod.f90:148-166), in one persistent-kernel call (bmx_track_bunch). Loss bookkeeping
! as bmx_track1_bunch_hook; survivors end with %ix_ele = ele2 (n_ele for ele2 = 0).
!-
subroutine bmx_track_bunch_span (lat, bunch, ele1, ele2, err, turn)
  type(bmx_lat_struct), intent(in) :: lat
  type(bmx_bunch_struct), intent(inout) :: bunch
  integer, intent(in) :: ele1, ele2
  logical, intent(out) :: err
  integer, optional, intent(in) :: turn
  integer(c_int32_t) :: t

  t = 0
  if (present(turn)) t = turn
  call bunch_through_device(lat, bunch, err, 2, int(ele1, c_int32_t), int(ele2, c_int32_t), t)
end subroutine

!+
! Subroutine bmx_bunch_free (bunch)
! Releases the device copy of the bunch (the particles stay).
!-
subroutine bmx_bunch_free (bunch)
  type(bmx_bunch_struct), intent(inout) :: bunch
  integer(c_int32_t) :: st
  if (c_associated(bunch%beam)) st = c_beam_destroy(bunch%beam)
  bunch%beam = c_null_ptr
  bunch%n_device = 0
end subroutine

! The bunch up to the device, one call (mode 1: element a of bmx_track_element; mode 2: span
! a -> b of bmx_track_bunch) with turn index t, and back.
subroutine bunch_through_device (lat, bunch, err, mode, a, b, t)
  type(bmx_lat_struct), intent(in) :: lat
  type(bmx_bunch_struct), intent(inout) :: bunch
  logical, intent(out) :: err
  integer, intent(in) :: mode
  integer(c_int32_t), intent(in) :: a, b, t

  real(c_double), allocatable, target :: x(:), px(:), y(:), py(:), z(:), pz(:), beta(:)
  integer(c_int32_t), allocatable, target :: state(:), loss_ele(:), loss_turn(:)
  type(bmx_beam_view) :: view
  integer(c_int32_t) :: st
  integer(c_int64_t) :: n
  integer :: i, ix_end

  err = .true.
  n = size(bunch%particle, kind = c_int64_t)
  if (n == 0) then
    err = .false.
    return
  endif
  allocate(x(n), px(n), y(n), py(n), z(n), pz(n), beta(n), state(n), loss_ele(n), loss_turn(n))
  do i = 1, int(n)
    x(i) = bunch%particle(i)%vec(1); px(i) = bunch%particle(i)%vec(2)
    y(i) = bunch%particle(i)%vec(3); py(i) = bunch%particle(i)%vec(4)
    z(i) = bunch%particle(i)%vec(5); pz(i) = bunch%particle(i)%vec(6)
    beta(i) = bunch%particle(i)%beta
    state(i) = bunch%particle(i)%state
    loss_ele(i) = 0
    loss_turn(i) = -1
  enddo
  call fill_view(view, x, px, y, py, z, pz, beta, state, loss_ele, loss_turn)
  if (bunch%n_device /= n) then
    call bmx_bunch_free(bunch)
    st = c_beam_create(lat%ctx, n, bunch%beam)
    if (st /= bmx_ok) return
    bunch%n_device = n
  endif
  st = c_beam_upload(bunch%beam, view, n)
  if (st == bmx_ok) then
    if (mode == 1) then
      st = c_track_element(lat%ctx, lat%lat, bunch%beam, a, t)
    else
      st = c_track_bunch(lat%ctx, lat%lat, bunch%beam, a, b, t)
    endif
  endif
  if (st == bmx_ok) st = c_beam_download(bunch%beam, view, n)
  if (st /= bmx_ok) return

  ix_end = a
  if (mode == 2) then
    ix_end = b
    if (ix_end == 0) ix_end = lat%n_ele
  endif
  do i = 1, int(n)
    if (bunch%particle(i)%state /= bmx_alive) cycle
    bunch%particle(i)%vec = [x(i), px(i), y(i), py(i), z(i), pz(i)]
    bunch%particle(i)%beta = beta(i)
    bunch%particle(i)%state = state(i)
    if (state(i) == bmx_alive) then
      bunch%particle(i)%ix_ele = ix_end
      bunch%particle(i)%ix_turn = t
    else
      bunch%particle(i)%ix_ele = loss_ele(i)
      bunch%particle(i)%ix_turn = loss_turn(i)
    endif
  enddo
  err = .false.
end subroutine

subroutine fill_view (view, x, px, y, py, z, pz, beta, state, loss_ele, loss_turn)
  type(bmx_beam_view), intent(out) :: view
  real(c_double), target, intent(in) :: x(:), px(:), y(:), py(:), z(:), pz(:), beta(:)
  integer(c_int32_t), target, intent(in) :: state(:), loss_ele(:), loss_turn(:)
  view%x = c_loc(x); view%px = c_loc(px); view%y = c_loc(y); view%py = c_loc(py)
  view%z = c_loc(z); view%pz = c_loc(pz); view%beta = c_loc(beta)
  view%state = c_loc(state); view%loss_ele = c_loc(loss_ele); view%loss_turn = c_loc(loss_turn)
end subroutine

end module
