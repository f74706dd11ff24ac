! Fortran caller of the regular-force library, in both binding styles:
! the ISO_C_BINDING module (gpunb_iface) and the implicit interface the
! reference host uses (CALL GPUNB_REGF(...), util_gpu.F:60).
! Input  (stream file): n, lmax, nnbmax, m_flag, m(n), x(3,n), v(3,n),
!                       h2(n), dtr(n)
!                       then the irregular-force block: lmax2, ni2, time,
!                       f(3,n), fdot(3,n), t0(n), list2(lmax2,n), addr(ni2)
! Output (stream file): acc(3,n), jrk(3,n), pot(n), list(lmax,n) from the
!                       module call, then the same from the implicit call,
!                       then pot(n) from gpupot, then firr(3,ni2),
!                       fd(3,ni2), iminr(ni2) from the irr_simd calls
!                       (legacy_irr.f, OpenMP set-up loop as intgrt.F).
program drv
  use iso_c_binding, only: c_int, c_double
  use gpunb_iface
  implicit none
  integer(c_int) :: n, lmax, nnbmax, mflag, irank, nbmax, one
  real(c_double), allocatable :: m(:), x(:,:), v(:,:), h2(:), dtr(:)
  real(c_double), allocatable :: acc(:,:), jrk(:,:), pot(:), phi(:)
  integer(c_int), allocatable :: list(:,:)
  integer(c_int) :: lmax2, ni2
  real(c_double) :: tirr
  real(c_double), allocatable :: f(:,:), fdot(:,:), t0(:), firr(:,:), fdirr(:,:)
  integer(c_int), allocatable :: list2(:,:), addr(:), iminr(:)
  character(len=512) :: fin, fout
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) n, lmax, nnbmax, mflag
  allocate(m(n), x(3,n), v(3,n), h2(n), dtr(n))
  allocate(acc(3,n), jrk(3,n), pot(n), phi(n), list(lmax,n))
  read(10) m, x, v, h2, dtr
  read(10) lmax2, ni2, tirr
  allocate(f(3,n), fdot(3,n), t0(n), list2(lmax2,n), addr(ni2))
  allocate(firr(3,ni2), fdirr(3,ni2), iminr(ni2))
  read(10) f, fdot, t0, list2, addr
  close(10)
  irank = 0
  nbmax = n + 10
  one = 1
  call gpunb_open(nbmax, irank)
  call gpunb_send(n, m, x, v)
  call gpunb_regf(n, h2, dtr, x, v, acc, jrk, pot, lmax, nnbmax, list, mflag)
  open(11, file=trim(fout), access='stream', form='unformatted', &
       status='replace')
  write(11) acc, jrk, pot, list
  acc = 0.0d0
  list = 0
  call legacy_regf(n, h2, dtr, x, v, acc, jrk, pot, lmax, nnbmax, list, mflag)
  write(11) acc, jrk, pot, list
  call gpunb_close()
  call gpupot(irank, one, n, n, m, x, phi)
  write(11) phi
  call legacy_irr(n, lmax2, x, v, f, fdot, m, t0, list2, tirr, ni2, addr, &
                  firr, fdirr, iminr)
  write(11) firr, fdirr, iminr
  close(11)
end program drv

