! mskg_bindings.f90 -- iso_c_binding interfaces to the MI355X KMC engine
! (include/mskg.h) for the Fortran host of MulSKIPS.
!
! The host keeps parsing start.dat and building the initial lattice
! (mulskips-source/main.f:25-300) and keeps writing frames (inpout/*.f); it
! replaces the KMC loop main.f:304-371 by calls into the engine:
!   mskg_create -> mskg_load_state -> { mskg_run -> mskg_store_state ->
!   WriteMolMolXYZFile } -> mskg_destroy.
! Scalars pass by reference as in mulskips-source-LaserAnnealing/fsockets.f
! (integer(c_int)), arrays as type(c_ptr), value (c_loc of the module arrays).
! Every function returns 0 on success; mskg_last_error gives the message.
module mskg_bindings
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: MSKG_OK = 0, MSKG_EINVAL = 1, MSKG_ESTATE = 2, &
                               MSKG_ECAPACITY = 3, MSKG_EDEVICE = 4, MSKG_ENOMEM = 5
  integer(c_int), parameter :: MSKG_STOP_LIMIT = 0, MSKG_STOP_FEW_MC = 1, MSKG_STOP_NO_EVENTS = 3
  integer(c_int32_t), parameter :: MSKG_FIELD_SOCKET = 0, MSKG_FIELD_PLANES = 1
  integer(c_int32_t), parameter :: MSKG_SCHEDULE_COLOURS = 0, MSKG_SCHEDULE_TILES = 1
  integer(c_int), parameter :: MSKG_KEY_ADS = 2496, MSKG_KEY_PH = 2499, MSKG_KEY_ISO = 2500, MSKG_KEY_N = 2501, &
                               MSKG_ADJ_NS = 2510

  ! mskg_config of include/mskg.h, field for field
  type, bind(c) :: mskg_config
    integer(c_int32_t) :: len_x, len_y, len_z
    integer(c_int32_t) :: n_crystal, n_cov
    integer(c_int32_t) :: list_crystal(3), list_cov(3)
    type(c_ptr) :: ptrans_e = c_null_ptr     ! c_loc(PTransE), 12288 doubles, Fortran order
    type(c_ptr) :: ptrans_d = c_null_ptr     ! c_loc(PTransD), 9 doubles
    type(c_ptr) :: ptrans_abs = c_null_ptr   ! c_loc(PTransAbs), 12 doubles
    type(c_ptr) :: ptrans_des = c_null_ptr   ! c_loc(PTransDes), 12 doubles
    real(c_double) :: ptrans_zig
    integer(c_int32_t) :: la_mode
    real(c_double) :: tm, p0, sigma, tflex
    integer(c_int64_t) :: seed
    integer(c_int32_t) :: tile, tile_capacity
    real(c_double) :: events_per_pass
    integer(c_int32_t) :: tau_policy, frozen_below_z
    type(c_ptr) :: second_factor = c_null_ptr
    integer(c_int32_t) :: stored_len_z = 0   ! planes held on the device (0: all)
    integer(c_int32_t) :: reserved(3)
  end type mskg_config

  ! one logged event (48 bytes)
  type, bind(c) :: mskg_event
    integer(c_int64_t) :: iter
    integer(c_int32_t) :: site(3)
    integer(c_int32_t) :: itrans, arm, coor
    real(c_double) :: rtot, time
  end type mskg_event

  interface
    integer(c_int) function mskg_create(cfg, eng) bind(c, name="mskg_create")
      import :: c_int, c_ptr, mskg_config
      type(mskg_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: eng
    end function

    subroutine mskg_destroy(eng) bind(c, name="mskg_destroy")
      import :: c_ptr
      type(c_ptr), value :: eng
    end subroutine

    type(c_ptr) function mskg_last_error(eng) bind(c, name="mskg_last_error")
      import :: c_ptr
      type(c_ptr), value :: eng
    end function

    ! LattCoo(0:LenX-1,0:LenY-1,0:LenZ) and the three lists: AtomXYZ as
    ! (3,n) int32, NextNXYZ as (3,4,n) int32, Ind_Event (n) int32,
    ! ProbTrans (n) real(8) -- pass c_loc of packed copies of the records.
    integer(c_int) function mskg_load_state(eng, latt_coo, n_ad, ad_xyz, ad_nn, ad_event, ad_prob, &
                                            n_bulk, bulk_xyz, bulk_nn, n_void, void_xyz, void_nn, &
                                            kmc_time) bind(c, name="mskg_load_state")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: eng, latt_coo, ad_xyz, ad_nn, ad_event, ad_prob
      integer(c_int32_t), value :: n_ad, n_bulk, n_void
      type(c_ptr), value :: bulk_xyz, bulk_nn, void_xyz, void_nn
      real(c_double), value :: kmc_time
    end function

    integer(c_int) function mskg_build_slab(eng, mode, len) bind(c, name="mskg_build_slab")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: eng
      integer(c_int32_t), value :: mode, len
    end function

    integer(c_int) function mskg_set_temperature(eng, temp_k) bind(c, name="mskg_set_temperature")
      import :: c_int, c_ptr
      type(c_ptr), value :: eng, temp_k
    end function

    ! field / coor / species: DIMENSION(LenZ, LenX*LenY) arrays a, b, c, d of
    ! the laser-annealing build with layout MSKG_FIELD_SOCKET (0)
    integer(c_int) function mskg_set_temperature_field(eng, field, layout, on_device) &
        bind(c, name="mskg_set_temperature_field")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: eng, field
      integer(c_int32_t), value :: layout, on_device
    end function

    integer(c_int) function mskg_phases(eng, coor, species, layout, on_device) bind(c, name="mskg_phases")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: eng, coor, species
      integer(c_int32_t), value :: layout, on_device
    end function

    integer(c_int) function mskg_set_la_variant(eng, p0, tm, sigma, tflex, init_count, count) &
        bind(c, name="mskg_set_la_variant")
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: eng
      real(c_double), intent(in) :: p0(3), tm(3), sigma(3), tflex(3)
      integer(c_int64_t), intent(in) :: init_count(3), count(3)
    end function

    ! pe(12288, ntemp), pd(9, ntemp), pabs(12, ntemp), pdes(12, ntemp): the
    ! refPtrans* arrays of SetProbabilityFromField.f with the table index last
    integer(c_int) function mskg_set_st_tables(eng, ntemp, temps, pe, pd, pabs, pdes) &
        bind(c, name="mskg_set_st_tables")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: eng, temps, pe, pd, pabs, pdes
      integer(c_int32_t), value :: ntemp
    end function

    integer(c_int) function mskg_species_counts(eng, evaporated, desorbed) bind(c, name="mskg_species_counts")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: eng
      integer(c_int64_t), intent(out) :: evaporated(3), desorbed(3)
    end function

    integer(c_int) function mskg_la_counts(eng, count, liquid) bind(c, name="mskg_la_counts")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: eng
      integer(c_int64_t), intent(out) :: count(3), liquid(3)
    end function

    integer(c_int) function mskg_refresh_rates(eng) bind(c, name="mskg_refresh_rates")
      import :: c_int, c_ptr
      type(c_ptr), value :: eng
    end function

    integer(c_int) function mskg_refresh_active(eng) bind(c, name="mskg_refresh_active")
      import :: c_int, c_ptr
      type(c_ptr), value :: eng
    end function

    integer(c_int) function mskg_run(eng, max_events, max_time, events_done, kmc_time, stop_reason) &
        bind(c, name="mskg_run")
      import :: c_int, c_int32_t, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: eng
      integer(c_int64_t), value :: max_events
      real(c_double), value :: max_time
      integer(c_int64_t), intent(out) :: events_done
      real(c_double), intent(out) :: kmc_time
      integer(c_int32_t), intent(out) :: stop_reason
    end function

    integer(c_int) function mskg_set_event_log(eng, capacity) bind(c, name="mskg_set_event_log")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: eng
      integer(c_int64_t), value :: capacity
    end function

    integer(c_int) function mskg_get_event_log(eng, out, cap, n) bind(c, name="mskg_get_event_log")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: eng, out
      integer(c_int64_t), value :: cap
      integer(c_int64_t), intent(out) :: n
    end function

    integer(c_int) function mskg_replay(eng, log, n, clock_mode) bind(c, name="mskg_replay")
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      type(c_ptr), value :: eng, log
      integer(c_int64_t), value :: n
      integer(c_int32_t), value :: clock_mode
    end function

    integer(c_int) function mskg_store_state(eng, latt_coo, cap_ad, n_ad, ad_xyz, ad_nn, ad_event, ad_prob, &
                                             cap_bulk, n_bulk, bulk_xyz, bulk_nn, cap_void, n_void, void_xyz, &
                                             void_nn) bind(c, name="mskg_store_state")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: eng, latt_coo
      integer(c_int32_t), value :: cap_ad, cap_bulk, cap_void
      integer(c_int32_t), intent(out) :: n_ad, n_bulk, n_void
      type(c_ptr), value :: ad_xyz, ad_nn, ad_event, ad_prob, bulk_xyz, bulk_nn, void_xyz, void_nn
    end function

    integer(c_int) function mskg_site_rates(eng, n, xyz, cls, rate) bind(c, name="mskg_site_rates")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: eng
      integer(c_int64_t), value :: n
      type(c_ptr), value :: xyz, cls, rate
    end function

    integer(c_int) function mskg_counters(eng, counts, kmc_time, n_adatom, n_atoms, n_voids) &
        bind(c, name="mskg_counters")
      import :: c_int, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: eng
      integer(c_int64_t), intent(out) :: counts(8)
      real(c_double), intent(out) :: kmc_time
      integer(c_int64_t), intent(out) :: n_adatom, n_atoms, n_voids
    end function

    integer(c_int) function mskg_build_alloy_slab(eng, len, fraction) bind(c, name="mskg_build_alloy_slab")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: eng
      integer(c_int32_t), value :: len
      real(c_double), value :: fraction
    end function

    ! serial BKL on n engines in one launch: engines, max_events, max_time are arrays of n
    integer(c_int) function mskg_run_serial(engines, n, max_events, max_time, events_done, kmc_time, stop_reason) &
        bind(c, name="mskg_run_serial")
      import :: c_int, c_int32_t, c_int64_t, c_ptr, c_double
      type(c_ptr), intent(in) :: engines(*)
      integer(c_int32_t), value :: n
      integer(c_int64_t), intent(in) :: max_events(*)
      real(c_double), intent(in) :: max_time(*)
      integer(c_int64_t), intent(out) :: events_done(*)
      real(c_double), intent(out) :: kmc_time(*)
      integer(c_int32_t), intent(out) :: stop_reason(*)
    end function

    ! the shaped initial states of main.f:224-294 (init_st NUL-terminated, lens = Len1..Len4)
    integer(c_int) function mskg_build_shape(eng, init_st, lens) bind(c, name="mskg_build_shape")
      import :: c_int, c_int32_t, c_ptr, c_char
      type(c_ptr), value :: eng
      character(kind=c_char), intent(in) :: init_st(*)
      integer(c_int32_t), intent(in) :: lens(4)
    end function

    ! SetCAD: the geometry ids (one int8 per grid point, x fastest, z = 0..LenZ-1)
    integer(c_int) function mskg_build_cad(eng, ids, on_device, alloy_fraction) bind(c, name="mskg_build_cad")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: eng
      type(c_ptr), value :: ids
      integer(c_int32_t), value :: on_device
      real(c_double), value :: alloy_fraction
    end function

    ! the batched synchronous-sublattice ensemble (same arguments as mskg_run_serial)
    integer(c_int) function mskg_run_ensemble(engines, n, max_events, max_time, events_done, kmc_time, &
        stop_reason) bind(c, name="mskg_run_ensemble")
      import :: c_int, c_int32_t, c_int64_t, c_ptr, c_double
      type(c_ptr), intent(in) :: engines(*)
      integer(c_int32_t), value :: n
      integer(c_int64_t), intent(in) :: max_events(*)
      real(c_double), intent(in) :: max_time(*)
      integer(c_int64_t), intent(out) :: events_done(*)
      real(c_double), intent(out) :: kmc_time(*)
      integer(c_int32_t), intent(out) :: stop_reason(*)
    end function

    ! trajectory-likelihood statistics: logs and dts are arrays of n pointers
    integer(c_int) function mskg_adjoint(engines, n, logs, dts, n_events, S, A, scalars) bind(c, name="mskg_adjoint")
      import :: c_int, c_int32_t, c_int64_t, c_ptr, c_double
      type(c_ptr), intent(in) :: engines(*), logs(*), dts(*)
      integer(c_int32_t), value :: n
      integer(c_int64_t), intent(in) :: n_events(*)
      real(c_double), intent(out) :: S(*), A(*), scalars(*)
    end function

    ! ... with class-population checkpoints every `every` events (hist: n pointers)
    integer(c_int) function mskg_adjoint_checkpointed(engines, n, logs, dts, n_events, S, A, scalars, every, hist) &
        bind(c, name="mskg_adjoint_checkpointed")
      import :: c_int, c_int32_t, c_int64_t, c_ptr, c_double
      type(c_ptr), intent(in) :: engines(*), logs(*), dts(*), hist(*)
      integer(c_int32_t), value :: n
      integer(c_int64_t), intent(in) :: n_events(*)
      real(c_double), intent(out) :: S(*), A(*), scalars(*)
      integer(c_int64_t), value :: every
    end function

    integer(c_int) function mskg_track_classes(eng, on) bind(c, name="mskg_track_classes")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: eng
      integer(c_int32_t), value :: on
    end function

    integer(c_int) function mskg_copy_state(dst, src) bind(c, name="mskg_copy_state")
      import :: c_int, c_ptr
      type(c_ptr), value :: dst, src
    end function

    ! frame compaction: (x,y,z,code) int32 records in grid order; counts(5)
    integer(c_int) function mskg_frame(eng, main_rec, cap_main, void_rec, cap_void, noepi_rec, cap_noepi, counts) &
        bind(c, name="mskg_frame")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: eng, main_rec, void_rec, noepi_rec
      integer(c_int64_t), value :: cap_main, cap_void, cap_noepi
      integer(c_int64_t), intent(out) :: counts(5)
    end function

    ! mskg_frame without an allocation: device buffers and a device counts(5)
    integer(c_int) function mskg_frame_device(eng, main_rec, cap_main, void_rec, cap_void, noepi_rec, cap_noepi, &
                                              counts) bind(c, name="mskg_frame_device")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: eng, main_rec, void_rec, noepi_rec, counts
      integer(c_int64_t), value :: cap_main, cap_void, cap_noepi
    end function

    ! the SiC defect frames (_v vacancy classes, _w wrong occupations); counts(2)
    integer(c_int) function mskg_frame_defects(eng, vac_rec, cap_vac, wrong_rec, cap_wrong, counts) &
        bind(c, name="mskg_frame_defects")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: eng, vac_rec, wrong_rec
      integer(c_int64_t), value :: cap_vac, cap_wrong
      integer(c_int64_t), intent(out) :: counts(2)
    end function

    integer(c_int) function mskg_last_timing(eng, run_ms, sweep_ms, sweep_points, sweeps_done) &
        bind(c, name="mskg_last_timing")
      import :: c_int, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: eng
      real(c_double), intent(out) :: run_ms, sweep_ms
      integer(c_int64_t), intent(out) :: sweep_points, sweeps_done
    end function

    integer(c_int) function mskg_run_diagnostics(eng, guard_exits, sweeps, tau, n_mc) &
        bind(c, name="mskg_run_diagnostics")
      import :: c_int, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: eng
      integer(c_int64_t), intent(out) :: guard_exits, sweeps, n_mc
      real(c_double), intent(out) :: tau
    end function

    ! MSKG_SCHEDULE_COLOURS (0, the default) or MSKG_SCHEDULE_TILES (1)
    integer(c_int) function mskg_set_schedule(eng, schedule) bind(c, name="mskg_set_schedule")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: eng
      integer(c_int32_t), value :: schedule
    end function
  end interface

contains

  ! mskg_last_error as a Fortran string
  function mskg_error_text(eng) result(msg)
    type(c_ptr), intent(in) :: eng
    character(len=512) :: msg
    character(kind=c_char), pointer :: cs(:)
    type(c_ptr) :: p
    integer :: i
    msg = ' '
    p = mskg_last_error(eng)
    if (.not. c_associated(p)) return
    call c_f_pointer(p, cs, [512])
    do i = 1, 512
      if (cs(i) == c_null_char) exit
      msg(i:i) = cs(i)
    end do
  end function

end module mskg_bindings
