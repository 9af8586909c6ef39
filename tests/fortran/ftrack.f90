!+
! Program ftrack
!
! A Fortran tracking host on the bmx_bmad_mod binding: reads a lattice and a bunch written by
! the test suite, tracks the bunch with bmx_track_bunch (track_bunch convention) and the first
! particle with bmx_track_all (track_all convention), and writes the results back.
!
! Usage: ftrack <dir>
!   <dir>/lattice.bin : int32 n_ele, real64 mc2, then n_ele bmx_ele records
!   <dir>/beam.bin    : int32 n, int32 n_turns, then x, px, y, py, z, pz, beta (n real64
!                       each) and state (n int32)
!   <dir>/out.bin     : x .. beta (n real64 each), state, ix_ele, ix_turn (n int32 each)
!   <dir>/orbit.bin   : int32 track_state, then for i = 0..n_ele: vec(6), beta, real(state)
!   <dir>/beam_out.dat: the tracked bunch as a Bmad beam file (bmx_write_beam_file)
!   <dir>/co.bin      : int32 err (0/1), 4-D closed orbit vec(6) from zero, one-turn matrix
!                       (6x6, row-major), with rel_tol 1e-12 and abs_tol 1e-15
!   <dir>/hook.bin    : the input bunch through one turn four ways -- element by element with
!                       bmx_track1_bunch_hook, whole with bmx_track_bunch_span(0, n_ele), and
!                       from element k = n_ele/3 round to k both ways (hooks k+1..n_ele, 1..k;
!                       bmx_track_bunch_span(k, k)) -- each as x .. beta (n real64 each), state,
!                       ix_ele, ix_turn (n int32 each), turn index 7
!-
program ftrack

use bmx_bmad_mod
use, intrinsic :: iso_c_binding
implicit none

type(bmx_ele_struct), allocatable :: eles(:)
type(bmx_coord_struct), allocatable :: p(:), orbit(:), p0(:)
type(bmx_bunch_struct) :: hb
logical :: finished
type(bmx_coord_struct) :: co
real(c_double) :: mat6(6, 6)
type(bmx_lat_struct) :: lat
integer(c_int32_t) :: n_ele, n, n_turns
integer(c_int32_t), allocatable :: ibuf(:)
real(c_double), allocatable :: buf(:)
real(c_double) :: mc2
integer :: u, i, k, track_state
logical :: err
character(1024) :: dir

call get_command_argument(1, dir)

open(newunit = u, file = trim(dir) // '/lattice.bin', access = 'stream', form = 'unformatted', status = 'old')
read(u) n_ele, mc2
allocate(eles(n_ele))
read(u) eles
close(u)

open(newunit = u, file = trim(dir) // '/beam.bin', access = 'stream', form = 'unformatted', status = 'old')
read(u) n, n_turns
allocate(p(n), buf(n), ibuf(n))
do k = 1, 6
  read(u) buf
  p(:)%vec(k) = buf
enddo
read(u) buf
p(:)%beta = buf
read(u) ibuf
p(:)%state = ibuf
close(u)
p(:)%p0c = eles(1)%p0c
p(:)%ix_turn = 0
p0 = p

call bmx_lat_init(lat, eles, mc2, -1, err)
if (err) then
  print '(a)', 'bmx_lat_init failed: ' // bmx_last_error_message()
  stop 1
endif

allocate(orbit(0:n_ele))
orbit(0) = p(1)

call bmx_track_bunch(lat, p, n_turns, err)
if (err) then
  print '(a)', 'bmx_track_bunch failed: ' // bmx_last_error_message()
  stop 2
endif

call bmx_write_beam_file(trim(dir) // '/beam_out.dat', p, 'Electron', 1e-12_c_double, int(n_ele), err)
if (err) then
  print '(a)', 'bmx_write_beam_file failed: ' // bmx_last_error_message()
  stop 4
endif

open(newunit = u, file = trim(dir) // '/out.bin', access = 'stream', form = 'unformatted', status = 'replace')
do k = 1, 6
  write(u) p(:)%vec(k)
enddo
write(u) p(:)%beta
write(u) int(p(:)%state, c_int32_t)
write(u) int(p(:)%ix_ele, c_int32_t)
write(u) int(p(:)%ix_turn, c_int32_t)
close(u)

call bmx_track_all(lat, orbit, track_state, err)
if (err .and. track_state == bmx_moving_forward) then
  print '(a)', 'bmx_track_all failed: ' // bmx_last_error_message()
  stop 3
endif
open(newunit = u, file = trim(dir) // '/orbit.bin', access = 'stream', form = 'unformatted', status = 'replace')
write(u) int(track_state, c_int32_t)
do i = 0, n_ele
  write(u) orbit(i)%vec, orbit(i)%beta, real(orbit(i)%state, c_double)
enddo
close(u)

co = bmx_coord_struct()
co%vec = 0
call bmx_closed_orbit_calc(lat, co, 4, err, mat6, 1e-12_c_double, 1e-15_c_double)
open(newunit = u, file = trim(dir) // '/co.bin', access = 'stream', form = 'unformatted', status = 'replace')
write(u) int(merge(1, 0, err), c_int32_t)
write(u) co%vec
write(u) transpose(mat6)
close(u)

open(newunit = u, file = trim(dir) // '/hook.bin', access = 'stream', form = 'unformatted', status = 'replace')
! one turn element by element through the hook
hb%particle = p0
do i = 1, n_ele
  call bmx_track1_bunch_hook(lat, hb, i, err, finished, turn = 7)
  if (err .or. .not. finished) then
    print '(a)', 'bmx_track1_bunch_hook failed: ' // bmx_last_error_message()
    stop 5
  endif
enddo
call write_bunch(u, hb%particle)
! the same turn in one span call
hb%particle = p0
call bmx_track_bunch_span(lat, hb, 0, int(n_ele), err, turn = 7)
if (err) stop 6
call write_bunch(u, hb%particle)
! a turn from the end of element k round to it: hooks, then one span call
k = int(n_ele) / 3
hb%particle = p0
do i = k + 1, n_ele
  call bmx_track1_bunch_hook(lat, hb, i, err, finished, turn = 7)
  if (err) stop 7
enddo
do i = 1, k
  call bmx_track1_bunch_hook(lat, hb, i, err, finished, turn = 7)
  if (err) stop 7
enddo
call write_bunch(u, hb%particle)
hb%particle = p0
call bmx_track_bunch_span(lat, hb, k, k, err, turn = 7)
if (err) stop 8
call write_bunch(u, hb%particle)
close(u)
call bmx_bunch_free(hb)

call bmx_lat_free(lat)
print '(a, i0, a, i0)', 'ftrack: tracked ', n, ' particles; track_state ', track_state

contains

subroutine write_bunch (u, q)
  integer, intent(in) :: u
  type(bmx_coord_struct), intent(in) :: q(:)
  integer :: kk
  do kk = 1, 6
    write(u) q(:)%vec(kk)
  enddo
  write(u) q(:)%beta
  write(u) int(q(:)%state, c_int32_t)
  write(u) int(q(:)%ix_ele, c_int32_t)
  write(u) int(q(:)%ix_turn, c_int32_t)
end subroutine

end program
