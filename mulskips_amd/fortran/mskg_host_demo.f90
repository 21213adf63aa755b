! mskg_host_demo.f90 -- the Fortran host's side of the engine, in miniature.
!
! The reference host (mulskips-source/main.f) builds its lattice, then runs
! the KMC loop of main.f:304-371.  With the engine the loop becomes
! mskg_load_state -> mskg_run (or mskg_replay of a recorded history) ->
! mskg_store_state, all through the iso_c_binding interfaces of
! mskg_bindings.f90.  This program does exactly that on a state handed over
! in a binary file (the host's module arrays, written by a test):
!
!   in:  int32  len_x len_y len_z n_crystal n_cov list_crystal(3) list_cov(3)
!               n_ad n_bulk n_void
!        real8  ptrans_zig PTransE(12288) PTransD(9) PTransAbs(12) PTransDes(12)
!        int64  n_events run_events
!        int32  LattCoo(len_x, len_y, len_z+1)
!        int32  AtomXYZ(3, n_ad) NextNXYZ(3, 4, n_ad) Ind_Event(n_ad); real8 ProbTrans(n_ad)
!        int32  AtomXYZ(3, n_bulk) NextNXYZ(3, 4, n_bulk)
!        int32  AtomXYZ(3, n_void) NextNXYZ(3, 4, n_void)
!        mskg_event log(n_events)
!   out: LattCoo after replaying the log, then the parallel run's event log
!        (run_events events of mskg_run) and LattCoo after it.
!
! usage: mskg_host_demo <in> <out>
program mskg_host_demo
  use iso_c_binding
  use mskg_bindings
  implicit none
  type(mskg_config), target :: cfg
  type(c_ptr) :: eng
  integer(c_int32_t) :: hdr(14), n_ad, n_bulk, n_void, n_got_ad, n_got_bulk, n_got_void, stop_reason
  integer(c_int64_t) :: n_events, run_events, done, n_log
  real(c_double), target :: zig, pe(12288), pd(9), pabs(12), pdes(12)
  real(c_double) :: kmc_time
  integer(c_int32_t), allocatable, target :: latt(:), ad_xyz(:), ad_nn(:), ad_ev(:), bulk_xyz(:), bulk_nn(:), &
                                             void_xyz(:), void_nn(:)
  real(c_double), allocatable, target :: ad_prob(:)
  type(mskg_event), allocatable, target :: log(:), run_log(:)
  character(len=4096) :: fin, fout
  integer(c_int64_t) :: npts

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) hdr
  read(10) zig, pe, pd, pabs, pdes
  cfg%len_x = hdr(1); cfg%len_y = hdr(2); cfg%len_z = hdr(3)
  cfg%n_crystal = hdr(4); cfg%n_cov = hdr(5)
  cfg%list_crystal = hdr(6:8); cfg%list_cov = hdr(9:11)
  n_ad = hdr(12); n_bulk = hdr(13); n_void = hdr(14)
  read(10) n_events, run_events
  npts = int(cfg%len_x, c_int64_t) * cfg%len_y * (cfg%len_z + 1)
  allocate(latt(npts), ad_xyz(3 * max(n_ad, 1)), ad_nn(12 * max(n_ad, 1)), ad_ev(max(n_ad, 1)), &
           ad_prob(max(n_ad, 1)), bulk_xyz(3 * max(n_bulk, 1)), bulk_nn(12 * max(n_bulk, 1)), &
           void_xyz(3 * max(n_void, 1)), void_nn(12 * max(n_void, 1)), log(max(n_events, 1_c_int64_t)), &
           run_log(max(run_events * 2, 1_c_int64_t)))
  read(10) latt
  if (n_ad > 0) read(10) ad_xyz(1:3 * n_ad), ad_nn(1:12 * n_ad), ad_ev(1:n_ad), ad_prob(1:n_ad)
  if (n_bulk > 0) read(10) bulk_xyz(1:3 * n_bulk), bulk_nn(1:12 * n_bulk)
  if (n_void > 0) read(10) void_xyz(1:3 * n_void), void_nn(1:12 * n_void)
  if (n_events > 0) read(10) log(1:n_events)
  close(10)

  cfg%ptrans_e = c_loc(pe); cfg%ptrans_d = c_loc(pd); cfg%ptrans_abs = c_loc(pabs); cfg%ptrans_des = c_loc(pdes)
  cfg%ptrans_zig = zig
  cfg%la_mode = 0; cfg%tm = 0; cfg%p0 = 0; cfg%sigma = 0; cfg%tflex = 0
  cfg%seed = 2026; cfg%tile = 24; cfg%tile_capacity = 0; cfg%events_per_pass = 1.0d0; cfg%tau_policy = 1
  cfg%frozen_below_z = 0; cfg%second_factor = c_null_ptr; cfg%stored_len_z = 0; cfg%reserved = 0
  call check(mskg_create(cfg, eng), 'mskg_create')
  call check(mskg_load_state(eng, c_loc(latt), n_ad, c_loc(ad_xyz), c_loc(ad_nn), c_loc(ad_ev), c_loc(ad_prob), &
                             n_bulk, c_loc(bulk_xyz), c_loc(bulk_nn), n_void, c_loc(void_xyz), c_loc(void_nn), &
                             0.0d0), 'mskg_load_state')
  ! the recorded history, then the state back into LattCoo
  call check(mskg_replay(eng, c_loc(log), n_events, 0_c_int32_t), 'mskg_replay')
  call check(mskg_store_state(eng, c_loc(latt), 0_c_int32_t, n_got_ad, c_null_ptr, c_null_ptr, c_null_ptr, &
                              c_null_ptr, 0_c_int32_t, n_got_bulk, c_null_ptr, c_null_ptr, 0_c_int32_t, &
                              n_got_void, c_null_ptr, c_null_ptr), 'mskg_store_state')
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) latt
  ! the KMC loop of main.f:304-371 as one engine call, logged
  call check(mskg_set_event_log(eng, run_events * 2), 'mskg_set_event_log')
  call check(mskg_run(eng, run_events, 1.0d300, done, kmc_time, stop_reason), 'mskg_run')
  call check(mskg_get_event_log(eng, c_loc(run_log), run_events * 2, n_log), 'mskg_get_event_log')
  call check(mskg_store_state(eng, c_loc(latt), 0_c_int32_t, n_got_ad, c_null_ptr, c_null_ptr, c_null_ptr, &
                              c_null_ptr, 0_c_int32_t, n_got_bulk, c_null_ptr, c_null_ptr, 0_c_int32_t, &
                              n_got_void, c_null_ptr, c_null_ptr), 'mskg_store_state')
  write(11) n_log, stop_reason, kmc_time
  write(11) run_log(1:n_log)
  write(11) latt
  close(11)
  call mskg_destroy(eng)
  write(*, '(A, I10, A, I10, A)') 'replayed ', n_events, ' events, ran ', n_log, ' events'

contains

  ! the reference STOPs on an inconsistent state; so does the host here
  subroutine check(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    if (code == MSKG_OK .or. (code == MSKG_ECAPACITY .and. what == 'mskg_store_state')) return
    write(*, '(A, A, I3, A, A)') what, ' failed: ', code, ' ', trim(mskg_error_text(eng))
    stop 1
  end subroutine

end program mskg_host_demo
