! ref_shim.f90 -- C entry points into the ParaGauss reference modules that
! compile standalone (oracle/ref.mk builds them, unmodified, from the
! reference tree).  TEST INFRASTRUCTURE ONLY: the resulting oracle/_ref
! library pins the oracle; nothing in the product links it.
!
! Wrapped reference routines:
!   pw_ldac(nv, iop=1 or 2, ...)       modules/pw_ldac_module.f90:198-270
!   pbe_ggcc("pbe", vla, ispin=1 or 2) modules/pbe_ggcxc_module.f90:214-349
!   solid_harmonics_calculate(_grads)  modules/solid_harmonics_module.f90:369-568
module pgref_shim
  use iso_c_binding, only: c_int, c_double
  implicit none
contains

  subroutine pgref_pw92(n, rho, f, fr) bind(C, name="pgref_pw92")
    use pw_ldac_module, only: pw_ldac
    integer(c_int), value :: n
    real(c_double), intent(in) :: rho(n)
    real(c_double), intent(out) :: f(n), fr(n)
    real(c_double) :: rr(n, 1), dr(n, 1)
    rr(:, 1) = rho
    f = 0
    dr = 0
    call pw_ldac(n, 1, rr, f, dr)
    fr = dr(:, 1)
  end subroutine pgref_pw92

  subroutine pgref_pbec(n, rho, gam, f, fr, fg) bind(C, name="pgref_pbec")
    use pbe_ggcxc_module, only: pbe_ggcc
    integer(c_int), value :: n
    real(c_double), intent(in) :: rho(n), gam(n)
    real(c_double), intent(out) :: f(n), fr(n), fg(n)
    real(c_double) :: rr(n, 1), gg(n, 1), dr(n, 1), dg(n, 1)
    rr(:, 1) = rho
    gg(:, 1) = gam
    f = 0
    dr = 0
    dg = 0
    call pbe_ggcc("pbe", n, 1, rr, gg, f, dr, dg)
    fr = dr(:, 1)
    fg = dg(:, 1)
  end subroutine pgref_pbec

  subroutine pgref_pw92_spin(n, na, nb, f, fa, fb) bind(C, name="pgref_pw92_spin")
    use pw_ldac_module, only: pw_ldac
    integer(c_int), value :: n
    real(c_double), intent(in) :: na(n), nb(n)
    real(c_double), intent(out) :: f(n), fa(n), fb(n)
    real(c_double) :: rr(n, 2), dr(n, 2)
    rr(:, 1) = na
    rr(:, 2) = nb
    f = 0
    dr = 0
    call pw_ldac(n, 2, rr, f, dr)
    fa = dr(:, 1)
    fb = dr(:, 2)
  end subroutine pgref_pw92_spin

  subroutine pgref_pbec_spin(n, na, nb, gaa, gbb, gab, f, fa, fb, fgaa, fgbb, fgab) &
       bind(C, name="pgref_pbec_spin")
    use pbe_ggcxc_module, only: pbe_ggcc
    integer(c_int), value :: n
    real(c_double), intent(in) :: na(n), nb(n), gaa(n), gbb(n), gab(n)
    real(c_double), intent(out) :: f(n), fa(n), fb(n), fgaa(n), fgbb(n), fgab(n)
    real(c_double) :: rr(n, 2), gg(n, 3), dr(n, 2), dg(n, 3)
    rr(:, 1) = na
    rr(:, 2) = nb
    gg(:, 1) = gaa
    gg(:, 2) = gbb
    gg(:, 3) = gab
    f = 0
    dr = 0
    dg = 0
    call pbe_ggcc("pbe", n, 2, rr, gg, f, dr, dg)
    fa = dr(:, 1)
    fb = dr(:, 2)
    fgaa = dg(:, 1)
    fgbb = dg(:, 2)
    fgab = dg(:, 3)
  end subroutine pgref_pbec_spin

  ! c(npts, (lmax+1)**2), g(npts, 3, (lmax+1)**2); pts(npts, 3) absolute
  subroutine pgref_solid_harmonics(npts, lmax, pts, center, c, g) &
       bind(C, name="pgref_solid_harmonics")
    use solid_harmonics_module
    use solhrules_module, only: solhrules_setup
    integer(c_int), value :: npts, lmax
    real(c_double), intent(in) :: pts(npts, 3), center(3)
    real(c_double), intent(out) :: c(npts, (lmax + 1)**2), g(npts, 3, (lmax + 1)**2)
    type(solid_harmonics_type) :: sh
    type(solid_harmonics_grads_type) :: shg
    integer :: l, m, lm
    logical, save :: rules_ready = .false.
    if (.not. rules_ready) then
       call solhrules_setup(6)
       rules_ready = .true.
    endif
    call solid_harmonics_setup(npts)
    call solid_harmonics_allocate(npts, lmax, sh=sh, shg=shg)
    call solid_harmonics_calculate(sh, pts, center, npts)
    call solid_harmonics_calculate_grads(sh, shg, npts)
    lm = 0
    do l = 0, lmax
       do m = 1, 2 * l + 1
          lm = lm + 1
          c(:, lm) = sh%l(l)%m(1:npts, m)
          g(:, :, lm) = shg%l(l)%m(1:npts, :, m)
       enddo
    enddo
    call solid_harmonics_free(sh=sh, shg=shg)
  end subroutine pgref_solid_harmonics
end module pgref_shim
