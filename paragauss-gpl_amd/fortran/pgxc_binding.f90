!
! pgxc_binding.f90 -- ISO_C_BINDING interface to include/pgxc.h.
!
! The Fortran SCF driver keeps its call interface: xc_build (ParaGauss
! modules/xc_hamiltonian.f90:1038) still calls calc_xcks_nl(ham_xc_arr)
! (:1130), whose body becomes a call of xc_gpu_build (xc_gpu.f90).  The
! enumerators below are the bit positions of xc_cntrl's bind(c) enum
! (modules/xc_cntrl.f90:61-104), so a functional mask is built from is_on().
!
module pgxc_binding
  use iso_c_binding, only: c_int, c_int32_t, c_int64_t, c_uint32_t, c_double, c_ptr, c_char
  implicit none
  private

  integer(c_int), parameter, public :: PGXC_OK = 0, PGXC_ERR_ARG = 1, &
       PGXC_ERR_UNSUPPORTED = 2, PGXC_ERR_HIP = 3, PGXC_ERR_NOMEM = 4, PGXC_ERR_STATE = 5

  ! xc_cntrl enumerators used by XC="PBE" (modules/xc_cntrl.f90:652-670)
  integer, parameter, public :: xc_XAlpha = 1, xc_PWLDAc = 3, xc_PBEx = 10, xc_PBEc = 18

  integer(c_int32_t), parameter, public :: PGXC_PREC_FP64 = 0, PGXC_PREC_FP32 = 1, &
       PGXC_PREC_BF16X3 = 2
  integer(c_int32_t), parameter, public :: PGXC_DENSITY_FULL = 0, PGXC_DENSITY_INCREMENTAL = 1

  public :: pgxc_create, pgxc_set_density, pgxc_set_orbitals, pgxc_build, pgxc_get_result
  public :: pgxc_energy, pgxc_gradient, pgxc_gradient_grid
  public :: pgxc_nbas, pgxc_last_error, pgxc_destroy, pgxc_last_timing, pgxc_last_stats
  public :: pgxc_last_work
  public :: pgxc_becke_weights, pgxc_create_opt, pgxc_layout, pgxc_options
  public :: pgxc_becke_weights_noalloc
  public :: pgxc_batch_create, pgxc_batch_set_priority, pgxc_batch_build, pgxc_batch_last_ms
  public :: pgxc_batch_last_error, pgxc_batch_destroy
  public :: pgxc_set_schedule, pgxc_batch_set_schedule, pgxc_set_lanes, pgxc_set_item_log
  public :: pgxc_state_bytes, pgxc_save_state, pgxc_load_state
  public :: pgxc_batch_set_item_log, pgxc_item_log, pgxc_batch_item_log
  public :: pgxc_costmodel_create, pgxc_costmodel_nparams, pgxc_costmodel_params
  public :: pgxc_costmodel_set_params, pgxc_costmodel_train, pgxc_costmodel_gradient
  public :: pgxc_costmodel_predict, pgxc_costmodel_last_error, pgxc_costmodel_destroy
  integer(c_int32_t), parameter, public :: PGXC_SCHED_STATIC = 0, PGXC_SCHED_ANALYTIC = 1, &
       PGXC_SCHED_MODEL = 2, PGXC_NFEATURES = 9

  type, bind(C) :: pgxc_options
     integer(c_int32_t) :: tile
     integer(c_int64_t) :: chunk_points
     integer(c_int32_t) :: variant_fp64
     integer(c_int32_t) :: variant_fp32
     integer(c_int32_t) :: no_screen
     integer(c_int32_t) :: no_basis_cache
  end type pgxc_options

  interface
     integer(c_int) function pgxc_create(ctx, n_atoms, atom_xyz, n_shells, shell_info, &
          exps, norms, contr, grid, npts, veclen) bind(C, name="pgxc_create")
       import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
       type(c_ptr), intent(out) :: ctx
       integer(c_int32_t), value :: n_atoms, n_shells, veclen
       real(c_double), intent(in) :: atom_xyz(*), exps(*), norms(*), contr(*), grid(*)
       integer(c_int32_t), intent(in) :: shell_info(*)
       integer(c_int64_t), value :: npts
     end function pgxc_create

     integer(c_int) function pgxc_set_density(ctx, ispin, nbas, P, mode) &
          bind(C, name="pgxc_set_density")
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ispin, nbas, mode
       real(c_double), intent(in) :: P(*)
     end function pgxc_set_density

     integer(c_int) function pgxc_set_orbitals(ctx, ispin, nbas, nocc, C, occ, mode) &
          bind(C, name="pgxc_set_orbitals")
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ispin, nbas, nocc, mode
       real(c_double), intent(in) :: C(*), occ(*)
     end function pgxc_set_orbitals

     integer(c_int) function pgxc_build(ctx, functional, precision, ham, exc, charge) &
          bind(C, name="pgxc_build")
       import :: c_int, c_int32_t, c_uint32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_uint32_t), value :: functional
       integer(c_int32_t), value :: precision
       real(c_double), intent(inout) :: ham(*), exc, charge
     end function pgxc_build

     integer(c_int) function pgxc_get_result(ctx, ham, exc, charge) bind(C, name="pgxc_get_result")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(inout) :: ham(*), exc, charge
     end function pgxc_get_result

     ! post-SCF XC energies per term (XAlpha, PWLDAc, PBEx, PBEc) and charge
     integer(c_int) function pgxc_energy(ctx, functional, precision, terms4, charge) &
          bind(C, name="pgxc_energy")
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: functional, precision
       real(c_double), intent(out) :: terms4(4), charge
     end function pgxc_energy

     ! dE_xc/dR, basis-centre term on the fixed grid: grad(3, n_atoms)
     integer(c_int) function pgxc_gradient(ctx, grad) bind(C, name="pgxc_gradient")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: grad(*)
     end function pgxc_gradient

     ! dE_xc/dR with the grid moving with the atoms (weight_grads on)
     integer(c_int) function pgxc_gradient_grid(ctx, owner, grad) bind(C, name="pgxc_gradient_grid")
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(in) :: owner(*)
       real(c_double), intent(out) :: grad(*)
     end function pgxc_gradient_grid

     integer(c_int64_t) function pgxc_nbas(ctx) bind(C, name="pgxc_nbas")
       import :: c_int64_t, c_ptr
       type(c_ptr), value :: ctx
     end function pgxc_nbas

     integer(c_int) function pgxc_last_timing(ctx, ms6) bind(C, name="pgxc_last_timing")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: ms6(6)
     end function pgxc_last_timing

     integer(c_int) function pgxc_last_stats(ctx, s5) bind(C, name="pgxc_last_stats")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: s5(5)
     end function pgxc_last_stats

     integer(c_int) function pgxc_last_work(ctx, w3) bind(C, name="pgxc_last_work")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: w3(3)
     end function pgxc_last_work

     type(c_ptr) function pgxc_last_error(ctx) bind(C, name="pgxc_last_error")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function pgxc_last_error

     integer(c_int) function pgxc_destroy(ctx) bind(C, name="pgxc_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function pgxc_destroy

     integer(c_int) function pgxc_becke_weights(n_atoms, atom_xyz, npts, pts, owner, &
          w_atomic, w_out) bind(C, name="pgxc_becke_weights")
       import :: c_int, c_int32_t, c_int64_t, c_double
       integer(c_int32_t), value :: n_atoms
       integer(c_int64_t), value :: npts
       real(c_double), intent(in) :: atom_xyz(*), pts(*), w_atomic(*)
       integer(c_int32_t), intent(in) :: owner(*)
       real(c_double), intent(out) :: w_out(*)
     end function pgxc_becke_weights

     ! the same weights from device arrays, with no device allocation
     ! (file-scope scratch table, at most 1024 atoms); queued on stream
     integer(c_int) function pgxc_becke_weights_noalloc(n_atoms, atom_xyz, npts, pts, owner, &
          w_atomic, w_out, stream) bind(C, name="pgxc_becke_weights_noalloc")
       import :: c_int, c_int32_t, c_int64_t, c_ptr
       integer(c_int32_t), value :: n_atoms
       integer(c_int64_t), value :: npts
       type(c_ptr), value :: atom_xyz, pts, owner, w_atomic, w_out, stream
     end function pgxc_becke_weights_noalloc

     integer(c_int) function pgxc_create_opt(ctx, n_atoms, atom_xyz, n_shells, shell_info, &
          exps, norms, contr, grid, npts, veclen, opt) bind(C, name="pgxc_create_opt")
       import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr, pgxc_options
       type(c_ptr), intent(out) :: ctx
       integer(c_int32_t), value :: n_atoms, n_shells, veclen
       real(c_double), intent(in) :: atom_xyz(*), exps(*), norms(*), contr(*), grid(*)
       integer(c_int32_t), intent(in) :: shell_info(*)
       integer(c_int64_t), value :: npts
       type(pgxc_options), intent(in) :: opt
     end function pgxc_create_opt

     integer(c_int) function pgxc_layout(ctx, out8) bind(C, name="pgxc_layout")
       import :: c_int, c_int64_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int64_t), intent(out) :: out8(8)
     end function pgxc_layout

     integer(c_int) function pgxc_batch_create(batch, n, ctxs) bind(C, name="pgxc_batch_create")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), intent(out) :: batch
       integer(c_int32_t), value :: n
       type(c_ptr), intent(in) :: ctxs(*)
     end function pgxc_batch_create

     integer(c_int) function pgxc_batch_set_priority(batch, cost) &
          bind(C, name="pgxc_batch_set_priority")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: batch
       real(c_double), intent(in) :: cost(*)
     end function pgxc_batch_set_priority

     integer(c_int) function pgxc_batch_build(batch, functional, precision, use_graph) &
          bind(C, name="pgxc_batch_build")
       import :: c_int, c_int32_t, c_uint32_t, c_ptr
       type(c_ptr), value :: batch
       integer(c_uint32_t), value :: functional
       integer(c_int32_t), value :: precision, use_graph
     end function pgxc_batch_build

     integer(c_int) function pgxc_batch_last_ms(batch, ms) bind(C, name="pgxc_batch_last_ms")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: batch
       real(c_double), intent(out) :: ms
     end function pgxc_batch_last_ms

     type(c_ptr) function pgxc_batch_last_error(batch) bind(C, name="pgxc_batch_last_error")
       import :: c_ptr
       type(c_ptr), value :: batch
     end function pgxc_batch_last_error

     integer(c_int) function pgxc_batch_destroy(batch) bind(C, name="pgxc_batch_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
     end function pgxc_batch_destroy
     integer(c_int) function pgxc_set_schedule(ctx, mode, model) bind(C, name="pgxc_set_schedule")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx, model
       integer(c_int32_t), value :: mode
     end function pgxc_set_schedule

     integer(c_int) function pgxc_batch_set_schedule(batch, mode, model) &
          bind(C, name="pgxc_batch_set_schedule")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: batch, model
       integer(c_int32_t), value :: mode
     end function pgxc_batch_set_schedule

     integer(c_int64_t) function pgxc_state_bytes(ctx) bind(C, name="pgxc_state_bytes")
       import :: c_int64_t, c_ptr
       type(c_ptr), value :: ctx
     end function pgxc_state_bytes

     integer(c_int) function pgxc_save_state(ctx, buf, bytes) bind(C, name="pgxc_save_state")
       import :: c_int, c_int64_t, c_ptr
       type(c_ptr), value :: ctx, buf
       integer(c_int64_t), value :: bytes
     end function pgxc_save_state

     integer(c_int) function pgxc_load_state(ctx, buf, bytes) bind(C, name="pgxc_load_state")
       import :: c_int, c_int64_t, c_ptr
       type(c_ptr), value :: ctx, buf
       integer(c_int64_t), value :: bytes
     end function pgxc_load_state

     integer(c_int) function pgxc_set_lanes(ctx, lanes) bind(C, name="pgxc_set_lanes")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: lanes
     end function pgxc_set_lanes

     integer(c_int) function pgxc_set_item_log(ctx, on) bind(C, name="pgxc_set_item_log")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: on
     end function pgxc_set_item_log

     integer(c_int) function pgxc_batch_set_item_log(batch, on) bind(C, name="pgxc_batch_set_item_log")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: batch
       integer(c_int32_t), value :: on
     end function pgxc_batch_set_item_log

     integer(c_int64_t) function pgxc_item_log(ctx, features, us, cap) bind(C, name="pgxc_item_log")
       import :: c_int64_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: features(*), us(*)
       integer(c_int64_t), value :: cap
     end function pgxc_item_log

     integer(c_int64_t) function pgxc_batch_item_log(batch, features, us, cap) &
          bind(C, name="pgxc_batch_item_log")
       import :: c_int64_t, c_double, c_ptr
       type(c_ptr), value :: batch
       real(c_double), intent(out) :: features(*), us(*)
       integer(c_int64_t), value :: cap
     end function pgxc_batch_item_log

     integer(c_int) function pgxc_costmodel_create(model, n_features, seed) &
          bind(C, name="pgxc_costmodel_create")
       import :: c_int, c_int32_t, c_int64_t, c_ptr
       type(c_ptr), intent(out) :: model
       integer(c_int32_t), value :: n_features
       integer(c_int64_t), value :: seed
     end function pgxc_costmodel_create

     integer(c_int32_t) function pgxc_costmodel_nparams(model) bind(C, name="pgxc_costmodel_nparams")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: model
     end function pgxc_costmodel_nparams

     integer(c_int) function pgxc_costmodel_params(model, params, norm) &
          bind(C, name="pgxc_costmodel_params")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: model
       real(c_double), intent(out) :: params(*), norm(*)
     end function pgxc_costmodel_params

     integer(c_int) function pgxc_costmodel_set_params(model, params, norm) &
          bind(C, name="pgxc_costmodel_set_params")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: model
       real(c_double), intent(in) :: params(*), norm(*)
     end function pgxc_costmodel_set_params

     integer(c_int) function pgxc_costmodel_train(model, x, y, n, steps, lr, loss) &
          bind(C, name="pgxc_costmodel_train")
       import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
       type(c_ptr), value :: model
       real(c_double), intent(in) :: x(*), y(*)
       integer(c_int64_t), value :: n
       integer(c_int32_t), value :: steps
       real(c_double), value :: lr
       real(c_double), intent(out) :: loss
     end function pgxc_costmodel_train

     integer(c_int) function pgxc_costmodel_gradient(model, x, y, n, grad, loss) &
          bind(C, name="pgxc_costmodel_gradient")
       import :: c_int, c_int64_t, c_double, c_ptr
       type(c_ptr), value :: model
       real(c_double), intent(in) :: x(*), y(*)
       integer(c_int64_t), value :: n
       real(c_double), intent(out) :: grad(*), loss
     end function pgxc_costmodel_gradient

     integer(c_int) function pgxc_costmodel_predict(model, x, n, out) &
          bind(C, name="pgxc_costmodel_predict")
       import :: c_int, c_int64_t, c_double, c_ptr
       type(c_ptr), value :: model
       real(c_double), intent(in) :: x(*)
       integer(c_int64_t), value :: n
       real(c_double), intent(out) :: out(*)
     end function pgxc_costmodel_predict

     type(c_ptr) function pgxc_costmodel_last_error(model) bind(C, name="pgxc_costmodel_last_error")
       import :: c_ptr
       type(c_ptr), value :: model
     end function pgxc_costmodel_last_error

     integer(c_int) function pgxc_costmodel_destroy(model) bind(C, name="pgxc_costmodel_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: model
     end function pgxc_costmodel_destroy
  end interface
end module pgxc_binding
