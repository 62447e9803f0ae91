!
! xc_gpu.f90 -- the Fortran side of the drop-in: what calc_xcks_nl /
! calc_xcks (ParaGauss modules/xc_hamiltonian.f90:1280-1580, :812-1036)
! become when their bodies call the MI355X library.
!
! Life cycle, mirroring the xc_hamiltonian module:
!   xc_setup    -> xc_gpu_setup(basis arrays, grdpts, veclen)      (once per geometry)
!   per SCF     -> xc_gpu_set_orbitals(eigvec_occ, occ_num_occ)    (density_calc input;
!                  rank 2 / 1 arrays: ispin = 1, rank 3 / 2: (.., ispin))
!   calc_xcks_nl-> xc_gpu_build(mask, precision, ham_xc_arr, exc_int, charge_int)
!   xc_close    -> xc_gpu_close()
! Non-zero library status is passed to xc_gpu_fail, which in ParaGauss is
! error_handler (error_handler.f90); the standalone driver stops.
!
module xc_gpu
  use iso_c_binding
  use pgxc_binding
  implicit none
  private

  public :: xc_gpu_setup, xc_gpu_set_orbitals, xc_gpu_set_density, xc_gpu_build, &
       xc_gpu_close, xc_gpu_mask, xc_gpu_status, xc_gpu_energy, xc_gpu_gradient, &
       xc_gpu_state_words, xc_gpu_load_state_words

  ! closed shell: eigvec_occ(nbas, nocc), occ_num_occ(nocc); open or closed
  ! shell as ParaGauss holds them: eigvec_occ(nbas, nocc, ispin),
  ! occ_num_occ(nocc, ispin)
  interface xc_gpu_set_orbitals
     module procedure set_orbitals_1, set_orbitals_s
  end interface xc_gpu_set_orbitals
  interface xc_gpu_set_density
     module procedure set_density_1, set_density_s
  end interface xc_gpu_set_density

  type(c_ptr), save :: ctx = c_null_ptr
  integer, save :: nbas_ctx = 0, ispin_ctx = 1
  integer(c_int), save :: last_status = PGXC_OK

contains

  integer function xc_gpu_status()
    xc_gpu_status = last_status
  end function xc_gpu_status

  subroutine xc_gpu_fail(where, rc)
    character(len=*), intent(in) :: where
    integer(c_int), intent(in) :: rc
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    last_status = rc
    text = ''
    if (c_associated(ctx)) then
       call c_f_pointer(pgxc_last_error(ctx), msg, [512])
       do i = 1, 512
          if (msg(i) == c_null_char) exit
          text(i:i) = msg(i)
       enddo
    endif
    write(*, '(a, a, a, i0, a, a)') 'xc_gpu: ', where, ': status ', rc, ': ', trim(text)
    error stop 1
  end subroutine xc_gpu_fail

  ! functional mask from the enabled xc_cntrl terms
  integer(c_uint32_t) function xc_gpu_mask(xalpha, pwldac, pbex, pbec)
    logical, intent(in) :: xalpha, pwldac, pbex, pbec
    xc_gpu_mask = 0
    if (xalpha) xc_gpu_mask = ior(xc_gpu_mask, ishft(1_c_int32_t, xc_XAlpha))
    if (pwldac) xc_gpu_mask = ior(xc_gpu_mask, ishft(1_c_int32_t, xc_PWLDAc))
    if (pbex) xc_gpu_mask = ior(xc_gpu_mask, ishft(1_c_int32_t, xc_PBEx))
    if (pbec) xc_gpu_mask = ior(xc_gpu_mask, ishft(1_c_int32_t, xc_PBEc))
  end function xc_gpu_mask

  subroutine xc_gpu_setup(atom_xyz, shell_info, exps, norms, contr, grdpts, veclen)
    real(c_double), intent(in) :: atom_xyz(:, :)        ! (3, n_atoms)
    integer(c_int32_t), intent(in) :: shell_info(:, :)  ! (5, n_shells)
    real(c_double), intent(in) :: exps(:), norms(:), contr(:)
    real(c_double), intent(in) :: grdpts(:, :)          ! (npts, 4) as grid_module
    integer, intent(in) :: veclen
    integer(c_int) :: rc
    call xc_gpu_close()
    rc = pgxc_create(ctx, int(size(atom_xyz, 2), c_int32_t), atom_xyz, &
         int(size(shell_info, 2), c_int32_t), shell_info, exps, norms, contr, grdpts, &
         int(size(grdpts, 1), c_int64_t), int(veclen, c_int32_t))
    if (rc /= PGXC_OK) call xc_gpu_fail('pgxc_create', rc)
    nbas_ctx = int(pgxc_nbas(ctx))
  end subroutine xc_gpu_setup

  integer(c_int32_t) function density_mode(incremental)
    logical, intent(in), optional :: incremental
    density_mode = PGXC_DENSITY_FULL
    if (present(incremental)) then
       if (incremental) density_mode = PGXC_DENSITY_INCREMENTAL
    endif
  end function density_mode

  ! eigvec_occ(nbas, nocc), occ_num_occ(nocc)  (modules/occupied_levels_module.f90:112-117)
  subroutine set_orbitals_1(eigvec_occ, occ_num_occ, incremental)
    real(c_double), intent(in) :: eigvec_occ(:, :), occ_num_occ(:)
    logical, intent(in), optional :: incremental
    integer(c_int) :: rc
    rc = pgxc_set_orbitals(ctx, 1_c_int32_t, int(size(eigvec_occ, 1), c_int32_t), &
         int(size(eigvec_occ, 2), c_int32_t), eigvec_occ, occ_num_occ, density_mode(incremental))
    if (rc /= PGXC_OK) call xc_gpu_fail('pgxc_set_orbitals', rc)
    ispin_ctx = 1
  end subroutine set_orbitals_1

  ! eigvec_occ(nbas, nocc, ispin), occ_num_occ(nocc, ispin)
  subroutine set_orbitals_s(eigvec_occ, occ_num_occ, incremental)
    real(c_double), intent(in) :: eigvec_occ(:, :, :), occ_num_occ(:, :)
    logical, intent(in), optional :: incremental
    integer(c_int) :: rc
    rc = pgxc_set_orbitals(ctx, int(size(eigvec_occ, 3), c_int32_t), &
         int(size(eigvec_occ, 1), c_int32_t), int(size(eigvec_occ, 2), c_int32_t), &
         eigvec_occ, occ_num_occ, density_mode(incremental))
    if (rc /= PGXC_OK) call xc_gpu_fail('pgxc_set_orbitals', rc)
    ispin_ctx = size(eigvec_occ, 3)
  end subroutine set_orbitals_s

  subroutine set_density_1(P, incremental)
    real(c_double), intent(in) :: P(:, :)
    logical, intent(in), optional :: incremental
    integer(c_int) :: rc
    rc = pgxc_set_density(ctx, 1_c_int32_t, int(size(P, 1), c_int32_t), P, density_mode(incremental))
    if (rc /= PGXC_OK) call xc_gpu_fail('pgxc_set_density', rc)
    ispin_ctx = 1
  end subroutine set_density_1

  ! P(nbas, nbas, ispin)
  subroutine set_density_s(P, incremental)
    real(c_double), intent(in) :: P(:, :, :)
    logical, intent(in), optional :: incremental
    integer(c_int) :: rc
    rc = pgxc_set_density(ctx, int(size(P, 3), c_int32_t), int(size(P, 1), c_int32_t), P, &
         density_mode(incremental))
    if (rc /= PGXC_OK) call xc_gpu_fail('pgxc_set_density', rc)
    ispin_ctx = size(P, 3)
  end subroutine set_density_s

  ! The body of calc_xcks_nl: accumulates into ham_xc_arr, exc_int,
  ! charge_int; ham_xc_arr has ispin * nbas (nbas + 1) / 2 entries, the spins
  ! interleaved as the reference stores them
  subroutine xc_gpu_build(mask, precision, ham_xc_arr, exc_int, charge_int)
    integer(c_uint32_t), intent(in) :: mask
    integer(c_int32_t), intent(in) :: precision
    real(c_double), intent(inout) :: ham_xc_arr(:)
    real(c_double), intent(inout) :: exc_int, charge_int
    integer(c_int) :: rc
    if (size(ham_xc_arr) /= ispin_ctx * nbas_ctx * (nbas_ctx + 1) / 2) &
         call xc_gpu_fail('xc_gpu_build: ham_xc_arr size', PGXC_ERR_ARG)
    rc = pgxc_build(ctx, mask, precision, ham_xc_arr, exc_int, charge_int)
    if (rc /= PGXC_OK) call xc_gpu_fail('pgxc_build', rc)
  end subroutine xc_gpu_build

  ! post_scf_calc_xc_energy: energy of each term (XAlpha, PWLDAc, PBEx, PBEc)
  ! on the current density, without the potential
  subroutine xc_gpu_energy(mask, precision, terms, charge)
    integer(c_uint32_t), intent(in) :: mask
    integer(c_int32_t), intent(in) :: precision
    real(c_double), intent(out) :: terms(4), charge
    integer(c_int) :: rc
    rc = pgxc_energy(ctx, int(mask, c_int32_t), precision, terms, charge)
    if (rc /= PGXC_OK) call xc_gpu_fail('pgxc_energy', rc)
  end subroutine xc_gpu_energy

  ! dE_xc/dR (3, n_atoms) at the last build: with owner (the atom of each
  ! grid point's atomic grid, 0-based) the grid moves with the atoms
  ! (weight_grads on), without it the grid is held (grad_xc = -grad)
  subroutine xc_gpu_gradient(grad, owner)
    real(c_double), intent(out) :: grad(:, :)
    integer(c_int32_t), intent(in), optional :: owner(:)
    integer(c_int) :: rc
    if (present(owner)) then
       rc = pgxc_gradient_grid(ctx, owner, grad)
       if (rc /= PGXC_OK) call xc_gpu_fail('pgxc_gradient_grid', rc)
    else
       rc = pgxc_gradient(ctx, grad)
       if (rc /= PGXC_OK) call xc_gpu_fail('pgxc_gradient', rc)
    endif
  end subroutine xc_gpu_gradient

  ! The library's incremental state as real(8) words for the SCF-state tape:
  ! words(1) = its size in bytes, then the bytes.  main_scf writes them with
  ! readwriteblocked_write right after xc_hamiltonian_store's records
  ! (modules/xc_hamiltonian.f90:2407-2486) and, on a restart, reads them back
  ! after xc_hamiltonian_recover (:2488-2634) and after xc_gpu_setup of the
  ! same geometry, so the next build can be incremental.
  subroutine xc_gpu_state_words(words)
    real(c_double), allocatable, target, intent(out) :: words(:)
    integer(c_int64_t) :: nb
    integer(c_int) :: rc
    nb = pgxc_state_bytes(ctx)
    if (nb < 0) call xc_gpu_fail('pgxc_state_bytes', int(-nb, c_int))
    allocate(words(1 + (nb + 7) / 8))
    words = 0.0_c_double
    words(1) = real(nb, c_double)
    rc = pgxc_save_state(ctx, c_loc(words(2)), nb)
    if (rc /= PGXC_OK) call xc_gpu_fail('pgxc_save_state', rc)
  end subroutine xc_gpu_state_words

  subroutine xc_gpu_load_state_words(words)
    real(c_double), target, intent(in) :: words(:)
    integer(c_int64_t) :: nb
    integer(c_int) :: rc
    nb = int(words(1), c_int64_t)
    if (size(words) < 1 + (nb + 7) / 8) call xc_gpu_fail('xc_gpu_load_state_words: short record', &
         PGXC_ERR_ARG)
    rc = pgxc_load_state(ctx, c_loc(words(2)), nb)
    if (rc /= PGXC_OK) call xc_gpu_fail('pgxc_load_state', rc)
  end subroutine xc_gpu_load_state_words

  subroutine xc_gpu_close()
    integer(c_int) :: rc
    if (c_associated(ctx)) then
       rc = pgxc_destroy(ctx)
       ctx = c_null_ptr
    endif
  end subroutine xc_gpu_close
end module xc_gpu
