! ndg_iface.f90 -- Fortran binding of the C ABI in include/ndyn_gpu.h (iso_c_binding).
!
! This is the interface a Fortran MD driver (Ndyn's src/nd.f90 'dynamics' or the
! driver in ndyn_gpu.f90) uses in place of compute0 / integrate1 / integrate2 /
! buildCellList / CheckListUpdate / mkIsoKE.  Arrays are passed in the reference's own
! layouts: pos(3,np), vel(3,np), typ(np), PairMat(Ntypes,Ntypes), tables (PTsize+1,NPairs).
module ndg_iface
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: NDG_OK = 0, NDG_ERR_NEIMAX = 1, NDG_ERR_OUTSIDE = 2, &
                               NDG_ERR_HIP = 3, NDG_ERR_ARG = 4, NDG_ERR_NOMEM = 5, &
                               NDG_ERR_RANGE = 6
  integer(c_int), parameter :: NDG_PATH_FP64 = 0, NDG_PATH_MIXED = 1
  integer(c_int), parameter :: NDG_TRIGGER_COMPAT = 0, NDG_TRIGGER_SKIN = 1
  integer(c_int), parameter :: NDG_MORSE = 1, NDG_BUCK = 2, NDG_LJ = 3, NDG_MORSE_PHYS = 4, &
                                NDG_SPRING = 5, NDG_SINE = 6, NDG_COSINE = 7, NDG_SOFT = 8, &
                                NDG_JOHNSON = 9, NDG_MORSE_C = 10, NDG_BUCK_C = 11, NDG_LJ_C = 12
  ! added to a kind: energy table V(r) - V(rc), force table unchanged (not in the reference)
  integer(c_int), parameter :: NDG_ESHIFT = 256
  integer(c_int), parameter :: NDG_FRAME_XYZ = 1, NDG_FRAME_DUMP = 2

  type, bind(C) :: ndg_run_params
    integer(c_int) :: nsteps
    real(c_double) :: dt
    integer(c_int) :: out_steps
    real(c_double) :: skin
    integer(c_int) :: renei, renei_freq
    integer(c_int) :: trigger
    integer(c_int) :: isokinetic
    real(c_double) :: ke_target, ke_tau
    integer(c_int) :: ke_freq
    integer(c_int) :: isobaric
    real(c_double) :: p_target(3), p_tau
    integer(c_int) :: p_freq
    integer(c_int) :: isoenergetic
    real(c_double) :: e_tau
    integer(c_int) :: e_freq
    integer(c_int) :: box_load
    real(c_double) :: box_vel(3)
  end type ndg_run_params

  type, bind(C) :: ndg_thermo
    integer(c_int) :: step
    integer(c_int) :: pad_
    real(c_double) :: pe, ke, eerr, pressure, volume
    real(c_double) :: virial(6)
    real(c_double) :: box(3)
  end type ndg_thermo

  interface
    integer(c_int) function ndg_pair_table_q(kind, params, qprd, ptrmin, ptsize, P, F, rc2, &
        dr2) bind(C, name='ndg_pair_table_q')
      import :: c_int, c_double
      integer(c_int), value :: kind, ptsize
      real(c_double), intent(in) :: params(4)
      real(c_double), value :: qprd, ptrmin
      real(c_double), intent(out) :: P(*), F(*), rc2, dr2
    end function

    integer(c_int) function ndg_pair_table(kind, params, ptrmin, ptsize, P, F, rc2, dr2) &
        bind(C, name='ndg_pair_table')
      import :: c_int, c_double
      integer(c_int), value :: kind, ptsize
      real(c_double), intent(in) :: params(4)
      real(c_double), value :: ptrmin
      real(c_double), intent(out) :: P(*), F(*), rc2, dr2
    end function

    integer(c_int) function ndg_create(h, np, ntypes, npairs, ptsize, neimax, flags) &
        bind(C, name='ndg_create')
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: h
      integer(c_int), value :: np, ntypes, npairs, ptsize, neimax, flags
    end function

    integer(c_int) function ndg_set_tables(h, P, F, rc2, dr2, ptrmin, pairmat, pairparams, &
        mass) bind(C, name='ndg_set_tables')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: P(*), F(*), rc2(*), dr2(*), pairparams(*), mass(*)
      real(c_double), value :: ptrmin
      integer(c_int), intent(in) :: pairmat(*)
    end function

    integer(c_int) function ndg_set_pair_kinds(h, kinds) bind(C, name='ndg_set_pair_kinds')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), intent(in) :: kinds(*)
    end function

    integer(c_int) function ndg_init_positions(h) bind(C, name='ndg_init_positions')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function

    integer(c_int) function ndg_set_initial_positions(h, ipos, ibox) &
        bind(C, name='ndg_set_initial_positions')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: ipos(*), ibox(3)
    end function

    integer(c_int) function ndg_set_holds(h, hold, nhold, bc_disp, bc_dir, bc_parm) &
        bind(C, name='ndg_set_holds')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: nhold
      integer(c_int), intent(in) :: hold(*), bc_disp(*), bc_dir(*)
      real(c_double), intent(in) :: bc_parm(*)
    end function

    integer(c_int) function ndg_set_state(h, pos, vel, acc, typ, box) &
        bind(C, name='ndg_set_state')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: pos(*), vel(*), acc(*), box(3)
      integer(c_int), intent(in) :: typ(*)
    end function

    integer(c_int) function ndg_get_state(h, pos, vel, acc, force, pene, box) &
        bind(C, name='ndg_get_state')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: pos(*), vel(*), acc(*), force(*), pene(*), box(3)
    end function

    integer(c_int) function ndg_build_lists(h, skin) bind(C, name='ndg_build_lists')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: skin
    end function

    integer(c_int) function ndg_integrate1(h, dt) bind(C, name='ndg_integrate1')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: dt
    end function

    integer(c_int) function ndg_compute(h, cv, cu, pe, virial) bind(C, name='ndg_compute')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: cv, cu
      real(c_double), intent(out) :: pe, virial(6)
    end function

    integer(c_int) function ndg_integrate2(h, dt, ck, ke) bind(C, name='ndg_integrate2')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: dt
      integer(c_int), value :: ck
      real(c_double), intent(out) :: ke
    end function

    integer(c_int) function ndg_scale_velocities(h, lambda) bind(C, name='ndg_scale_velocities')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: lambda
    end function

    integer(c_int) function ndg_check_rebuild(h, skin, trigger, rebuilt) &
        bind(C, name='ndg_check_rebuild')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: skin
      integer(c_int), value :: trigger
      integer(c_int), intent(out) :: rebuilt
    end function

    integer(c_int) function ndg_run_begin(h, rp, step0) bind(C, name='ndg_run_begin')
      import :: c_int, c_ptr, ndg_run_params
      type(c_ptr), value :: h
      type(ndg_run_params), intent(in) :: rp
      integer(c_int), value :: step0
    end function

    integer(c_int) function ndg_run_steps(h, nsteps, rec, max_rec, nrec) &
        bind(C, name='ndg_run_steps')
      import :: c_int, c_ptr, ndg_thermo
      type(c_ptr), value :: h
      integer(c_int), value :: nsteps, max_rec
      type(ndg_thermo), intent(out) :: rec(*)
      integer(c_int), intent(out) :: nrec
    end function

    integer(c_int) function ndg_run_enqueue(h, nsteps) bind(C, name='ndg_run_enqueue')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: nsteps
    end function

    integer(c_int) function ndg_run_wait(h, rec, max_rec, nrec) bind(C, name='ndg_run_wait')
      import :: c_int, c_ptr, ndg_thermo
      type(c_ptr), value :: h
      integer(c_int), value :: max_rec
      type(ndg_thermo), intent(out) :: rec(*)
      integer(c_int), intent(out) :: nrec
    end function

    integer(c_int) function ndg_snapshot(h) bind(C, name='ndg_snapshot')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function

    integer(c_int) function ndg_snapshot_get(h, pos, vel, force, pene, box, step) &
        bind(C, name='ndg_snapshot_get')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: pos(*), vel(*), force(*), pene(*), box(3)
      integer(c_int), intent(out) :: step
    end function

    integer(c_int) function ndg_frame_write(h, path, format, names, name_len) &
        bind(C, name='ndg_frame_write')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h
      character(kind=c_char), intent(in) :: path(*), names(*)
      integer(c_int), value :: format, name_len
    end function

    integer(c_int) function ndg_frame_flush(h) bind(C, name='ndg_frame_flush')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function

    integer(c_int) function ndg_format_frame(format, np, pos, vel, pene, typ, names, name_len, &
        box, step, buf, cap, len) bind(C, name='ndg_format_frame')
      import :: c_int, c_double, c_char, c_size_t
      integer(c_int), value :: format, np, name_len, step
      real(c_double), intent(in) :: pos(*), vel(*), pene(*), box(3)
      integer(c_int), intent(in) :: typ(*)
      character(kind=c_char), intent(in) :: names(*)
      character(kind=c_char), intent(out) :: buf(*)
      integer(c_size_t), value :: cap
      integer(c_size_t), intent(out) :: len
    end function

    integer(c_int) function ndg_run_end(h, nupdates, step) bind(C, name='ndg_run_end')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), intent(out) :: nupdates, step
    end function

    integer(c_int) function ndg_destroy(h) bind(C, name='ndg_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function

    type(c_ptr) function ndg_last_error() bind(C, name='ndg_last_error')
      import :: c_ptr
    end function

    integer(c_int) function ndg_set_box(h, box) bind(C, name='ndg_set_box')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double), intent(in) :: box(3)
    end function

    integer(c_int) function ndg_get_box(h, box) bind(C, name='ndg_get_box')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double), intent(out) :: box(3)
    end function

    ! neighbour list in caller order: counts(np), and nbr(cap, np) when present
    integer(c_int) function ndg_get_neighbours(h, counts, nbr, cap) &
        bind(C, name='ndg_get_neighbours')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), intent(out) :: counts(*)
      integer(c_int), intent(out), optional :: nbr(*)
      integer(c_int), value :: cap
    end function

    ! ---------------------------------------------------------------- batched engine
    ! many independent systems advanced per launch (see include/ndyn_gpu.h)

    integer(c_int) function ndg_engine_create(e, path, hbm_budget) &
        bind(C, name='ndg_engine_create')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), intent(out) :: e
      integer(c_int), value :: path
      integer(c_size_t), value :: hbm_budget
    end function

    integer(c_int) function ndg_engine_destroy(e) bind(C, name='ndg_engine_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
    end function

    integer(c_size_t) function ndg_footprint(np, ntypes, npairs, ptsize, neimax) &
        bind(C, name='ndg_footprint')
      import :: c_int, c_size_t
      integer(c_int), value :: np, ntypes, npairs, ptsize, neimax
    end function

    integer(c_int) function ndg_engine_admit(e, slot, np, ntypes, npairs, ptsize, neimax, &
        P, F, rc2, dr2, ptrmin, pairmat, pairparams, mass, pos, vel, typ, box, rp) &
        bind(C, name='ndg_engine_admit')
      import :: c_int, c_double, c_ptr, ndg_run_params
      type(c_ptr), value :: e
      integer(c_int), intent(out) :: slot
      integer(c_int), value :: np, ntypes, npairs, ptsize, neimax
      real(c_double), intent(in) :: P(*), F(*), rc2(*), dr2(*)
      real(c_double), value :: ptrmin
      integer(c_int), intent(in) :: pairmat(*), typ(*)
      real(c_double), intent(in) :: pairparams(*), mass(*), pos(*), vel(*), box(3)
      type(ndg_run_params), intent(in) :: rp
    end function

    ! kinds, ipos, hold, bc_disp, bc_dir, bc_parm: c_loc of the arrays or c_null_ptr (no
    ! pair kinds / take the admitted state as initial positions / no holds, nhold = 0)
    integer(c_int) function ndg_engine_admit_ex(e, slot, np, ntypes, npairs, ptsize, &
        neimax, P, F, rc2, dr2, ptrmin, pairmat, pairparams, mass, pos, vel, typ, box, rp, &
        kinds, ipos, init_box, hold, nhold, bc_disp, bc_dir, bc_parm) &
        bind(C, name='ndg_engine_admit_ex')
      import :: c_int, c_double, c_ptr, ndg_run_params
      type(c_ptr), value :: e
      integer(c_int), intent(out) :: slot
      integer(c_int), value :: np, ntypes, npairs, ptsize, neimax, nhold
      real(c_double), intent(in) :: P(*), F(*), rc2(*), dr2(*)
      real(c_double), value :: ptrmin
      integer(c_int), intent(in) :: pairmat(*), typ(*)
      real(c_double), intent(in) :: pairparams(*), mass(*), pos(*), vel(*), box(3)
      type(ndg_run_params), intent(in) :: rp
      type(c_ptr), value :: kinds, ipos, hold, bc_disp, bc_dir, bc_parm
      real(c_double), intent(in) :: init_box(3)
    end function

    integer(c_int) function ndg_engine_step(e, nsteps, ndone) bind(C, name='ndg_engine_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: nsteps
      integer(c_int), intent(out) :: ndone
    end function

    integer(c_int) function ndg_engine_status(e, slot, steps_done, finished, last, err) &
        bind(C, name='ndg_engine_status')
      import :: c_int, c_ptr, ndg_thermo
      type(c_ptr), value :: e
      integer(c_int), value :: slot
      integer(c_int), intent(out) :: steps_done, finished, err
      type(ndg_thermo), intent(out) :: last
    end function

    integer(c_int) function ndg_engine_thermo(e, slot, rec, max_rec, nrec) &
        bind(C, name='ndg_engine_thermo')
      import :: c_int, c_ptr, ndg_thermo
      type(c_ptr), value :: e
      integer(c_int), value :: slot, max_rec
      type(ndg_thermo), intent(out) :: rec(*)
      integer(c_int), intent(out) :: nrec
    end function

    integer(c_int) function ndg_engine_fetch(e, slot, pos, vel, force, box) &
        bind(C, name='ndg_engine_fetch')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: slot
      real(c_double), intent(out) :: pos(*), vel(*), force(*), box(3)
    end function

    integer(c_int) function ndg_engine_step_enqueue(e, nsteps) &
        bind(C, name='ndg_engine_step_enqueue')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: nsteps
    end function

    integer(c_int) function ndg_engine_step_wait(e, ndone) bind(C, name='ndg_engine_step_wait')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), intent(out) :: ndone
    end function

    integer(c_int) function ndg_engine_cell_blocks(e, n, np_max, by, bz) &
        bind(C, name='ndg_engine_cell_blocks')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: n
      integer(c_int), intent(in) :: np_max(*), by(*), bz(*)
    end function

    ! ndg_engine_fetch without allocation (systems of at most 262144 atoms)
    integer(c_int) function ndg_engine_fetch_noalloc(e, slot, pos, vel, force, box) &
        bind(C, name='ndg_engine_fetch_noalloc')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: slot
      real(c_double), intent(out) :: pos(*), vel(*), force(*), box(3)
    end function

    integer(c_int) function ndg_engine_release(e, slot) bind(C, name='ndg_engine_release')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: slot
    end function

    integer(c_int) function ndg_engine_config(e, tpa, graph_steps) &
        bind(C, name='ndg_engine_config')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: tpa, graph_steps
    end function

    integer(c_int) function ndg_engine_size_classes(e, n, np_max, tpa) &
        bind(C, name='ndg_engine_size_classes')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: n
      integer(c_int), intent(in) :: np_max(*), tpa(*)
    end function

    integer(c_int) function ndg_engine_kernel_time(e, reset, avg_ms, count, pairs, atoms) &
        bind(C, name='ndg_engine_kernel_time')
      import :: c_int, c_double, c_long_long, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: reset
      real(c_double), intent(out) :: avg_ms, pairs, atoms
      integer(c_long_long), intent(out) :: count
    end function

    integer(c_int) function ndg_engine_timing(e, enable) bind(C, name='ndg_engine_timing')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: enable
    end function

    integer(c_int) function ndg_engine_counters(e, slot, nupdates, step, mean_nbr) &
        bind(C, name='ndg_engine_counters')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: slot
      integer(c_int), intent(out) :: nupdates, step
      real(c_double), intent(out) :: mean_nbr
    end function

    integer(c_int) function ndg_engine_list_stats(e, mean_nbr, atoms) &
        bind(C, name='ndg_engine_list_stats')
      import :: c_int, c_double, c_long_long, c_ptr
      type(c_ptr), value :: e
      real(c_double), intent(out) :: mean_nbr
      integer(c_long_long), intent(out) :: atoms
    end function

    integer(c_int) function ndg_copy_bandwidth(bytes, iters, gbps) &
        bind(C, name='ndg_copy_bandwidth')
      import :: c_int, c_double, c_size_t
      integer(c_size_t), value :: bytes
      integer(c_int), value :: iters
      real(c_double), intent(out) :: gbps
    end function
  end interface

contains

  ! message of the last failing call as a Fortran string
  function ndg_error_text() result(msg)
    character(len=256) :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: cs(:)
    integer :: k
    msg = ' '
    p = ndg_last_error()
    if (.not. c_associated(p)) return
    call c_f_pointer(p, cs, [256])
    do k = 1, 256
      if (cs(k) == c_null_char) exit
      msg(k:k) = cs(k)
    enddo
  end function ndg_error_text

end module ndg_iface
