!
! pgxc_driver -- standalone Fortran caller of the library, in the shape the
! ParaGauss SCF driver calls it (setup once, then orbitals + build).
!
! usage: pgxc_driver <input.bin> <output.bin>
! input  (stream, native endianness): int32 n_atoms, n_shells, nbas, nocc,
!        veclen, precision, mask, ispin; int64 npts; real64 atom_xyz(3,n_atoms);
!        int32 shell_info(5,n_shells); int32 n_exp_total, n_contr_total;
!        real64 exps(n_exp_total), norms(n_exp_total), contr(n_contr_total),
!        grdpts(npts,4), eigvec_occ(nbas,nocc,ispin), occ_num_occ(nocc,ispin)
! output: real64 exc_int, charge_int, ham_xc_arr(ispin*nbas*(nbas+1)/2), then
!         the post-SCF energy terms(4) and charge of the same density, the
!         basis-centre gradient dE_xc/dR(3, n_atoms), and exc_int, charge_int,
!         ham_xc_arr of an incremental build of the same density after a
!         restart (state saved as tape words, a new setup, the words loaded)
!
program pgxc_driver
  use iso_c_binding
  use xc_gpu
  implicit none
  character(len=1024) :: fin, fout
  integer(c_int32_t) :: n_atoms, n_shells, nbas, nocc, veclen, precision, ispin, ne, nc
  integer(c_uint32_t) :: mask
  integer(c_int64_t) :: npts
  real(c_double), allocatable :: atom_xyz(:, :), exps(:), norms(:), contr(:), grdpts(:, :)
  real(c_double), allocatable :: eigvec_occ(:, :, :), occ(:, :), ham(:)
  integer(c_int32_t), allocatable :: shell_info(:, :)
  real(c_double) :: exc_int, charge_int, terms(4), charge_e
  real(c_double), allocatable :: grad(:, :), words(:), ham2(:)
  real(c_double) :: exc2, charge2
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) n_atoms, n_shells, nbas, nocc, veclen, precision, mask, ispin
  read(u) npts
  allocate(atom_xyz(3, n_atoms), shell_info(5, n_shells))
  read(u) atom_xyz
  read(u) shell_info
  read(u) ne, nc
  allocate(exps(ne), norms(ne), contr(max(nc, 1)), grdpts(npts, 4), &
       eigvec_occ(nbas, nocc, ispin), occ(nocc, ispin))
  read(u) exps, norms
  if (nc > 0) read(u) contr(1:nc)
  read(u) grdpts, eigvec_occ, occ
  close(u)

  ! xc_setup
  call xc_gpu_setup(atom_xyz, shell_info, exps, norms, contr, grdpts, int(veclen))
  ! xc_build: zero the accumulators (modules/xc_hamiltonian.f90:1101-1116),
  ! then the calc_xcks_nl body
  allocate(ham(ispin * nbas * (nbas + 1) / 2))
  ham = 0.0_c_double
  exc_int = 0.0_c_double
  charge_int = 0.0_c_double
  call xc_gpu_set_orbitals(eigvec_occ, occ)
  call xc_gpu_build(mask, precision, ham, exc_int, charge_int)
  ! post-SCF: energies per term and the gradient at this density
  allocate(grad(3, n_atoms))
  call xc_gpu_energy(mask, precision, terms, charge_e)
  call xc_gpu_gradient(grad)
  ! checkpoint / restart: the state beside the SCF-state records, a new
  ! setup of the same geometry, then an incremental build (dP = 0 here)
  call xc_gpu_state_words(words)
  call xc_gpu_close()
  call xc_gpu_setup(atom_xyz, shell_info, exps, norms, contr, grdpts, int(veclen))
  call xc_gpu_load_state_words(words)
  allocate(ham2(size(ham)))
  ham2 = 0.0_c_double
  exc2 = 0.0_c_double
  charge2 = 0.0_c_double
  call xc_gpu_set_orbitals(eigvec_occ, occ, incremental=.true.)
  call xc_gpu_build(mask, precision, ham2, exc2, charge2)
  call xc_gpu_close()

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) exc_int, charge_int, ham, terms, charge_e, grad, exc2, charge2, ham2
  close(u)
  write(*, '(a, es24.16, a, es24.16)') 'exc_int = ', exc_int, '  charge_int = ', charge_int
end program pgxc_driver
