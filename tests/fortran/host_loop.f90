! host_loop.f90 -- the reference host's regular-force call sequence, written
! against the library's implicit interfaces (no module), to exercise the
! drop-in the way src/Main/util_gpu.F:25-112 (CALC_REG_ON_GPU) drives it:
!
!   * the active list IREG is cut into chunks of MAXTHR = 1024 i
!     (util_gpu.F:7,36-37); the per-chunk working arrays are filled in an
!     OpenMP loop (util_gpu.F:41-55): h2 = RS^2 (RS^2 / BODYM with the mass
!     criterion), dtr = STEPR, positions and velocities;
!   * after every GPUNB_REGF the rows are scanned in order; the first row
!     with a negative count (overflow) shrinks that body's RS -- by
!     (NNBOPT/NNB)^0.333 or (NNBOPT/NNBMAX)^0.4 for smooth systems
!     (NB_FLAG = 1), by 0.9 otherwise -- and the whole chunk is computed
!     again (util_gpu.F:66-101);
!   * accepted rows get IFIRST added to every 0-based j, the body itself
!     dropped and the count rewritten; the potential goes to PHIDBL
!     (util_gpu.F:102-112).
!
! The j set sent is bodies IFIRST..NTOT in order (intgrt.F:912-918).
! Input  (stream): ntot, ifirst, lmax, nnbmax, nnbopt, m_flag, nb_flag,
!                  nreg, bodym, body(ntot), x(3,ntot), xdot(3,ntot),
!                  rs(ntot), stepr(ntot), ireg(nreg)
! Output (stream): rs(ntot) (after the shrinks), phidbl(ntot),
!                  freg(3,ntot), fdr(3,ntot), list(lmax,ntot) (rows of the
!                  active bodies; others zero), nretry
program host_loop
  implicit none
  integer, parameter :: maxthr = 1024
  integer :: ntot, ifirst, lmax, nnbmax, nnbopt, m_flag, nb_flag, nreg
  integer :: l, ni, ii, idi, nnb, ll, l1, itemp, nofl, nretry, irank, nbmax
  integer :: nj
  double precision :: bodym
  double precision, allocatable :: body(:), x(:,:), xdot(:,:), rs(:)
  double precision, allocatable :: stepr(:), phidbl(:), freg(:,:), fdr(:,:)
  integer, allocatable :: ireg(:), list(:,:), listgp(:,:)
  double precision :: gpu_rs(maxthr), gpu_dtr(maxthr), gpu_pot(maxthr)
  double precision :: x_i(3,maxthr), v_i(3,maxthr)
  double precision :: gpu_a(3,maxthr), gpu_jerk(3,maxthr)
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) ntot, ifirst, lmax, nnbmax, nnbopt, m_flag, nb_flag, nreg
  read(10) bodym
  allocate(body(ntot), x(3,ntot), xdot(3,ntot), rs(ntot), stepr(ntot))
  allocate(phidbl(ntot), freg(3,ntot), fdr(3,ntot), ireg(nreg))
  allocate(list(lmax,ntot), listgp(lmax,maxthr))
  read(10) body, x, xdot, rs, stepr, ireg
  close(10)
  phidbl = 0.0d0
  freg = 0.0d0
  fdr = 0.0d0
  list = 0
  nretry = 0

  irank = 0
  nbmax = ntot + 10
  nj = ntot - ifirst + 1
  call gpunb_open(nbmax, irank)
  call gpunb_send(nj, body(ifirst), x(1,ifirst), xdot(1,ifirst))

  do l = 1, nreg, maxthr
    ni = min(maxthr, nreg - l + 1)
!$omp parallel do private(ii, idi)
    do ii = 1, ni
      idi = ireg(l + ii - 1)
      if (m_flag == 1) then
        gpu_rs(ii) = rs(idi)**2 / bodym
      else
        gpu_rs(ii) = rs(idi)**2
      end if
      gpu_dtr(ii) = stepr(idi)
      x_i(1:3, ii) = x(1:3, idi)
      v_i(1:3, ii) = xdot(1:3, idi)
    end do
!$omp end parallel do
    do
      call gpunb_regf(ni, gpu_rs, gpu_dtr, x_i, v_i, gpu_a, gpu_jerk, &
                      gpu_pot, lmax, nnbmax, listgp, m_flag)
      nofl = 0
      do ii = 1, ni
        idi = ireg(l + ii - 1)
        nnb = listgp(1, ii)
        if (nnb < 0) then
          ! Shrink the first overflowing body's radius and redo the chunk.
          if (nb_flag == 1) then
            if (-nnb > nnbopt) then
              rs(idi) = (-real(nnbopt) / real(nnb))**0.333 * rs(idi)
            else
              rs(idi) = (real(nnbopt) / real(nnbmax))**0.4 * rs(idi)
            end if
          else
            rs(idi) = 0.9 * rs(idi)
          end if
          if (m_flag == 1) then
            gpu_rs(ii) = rs(idi)**2 / bodym
          else
            gpu_rs(ii) = rs(idi)**2
          end if
          nofl = 1
          exit
        end if
      end do
      if (nofl == 0) exit
      nretry = nretry + 1
    end do
    do ii = 1, ni
      idi = ireg(l + ii - 1)
      nnb = listgp(1, ii)
      l1 = 1
      do ll = 2, nnb + 1
        itemp = listgp(ll, ii) + ifirst
        if (itemp /= idi) then
          l1 = l1 + 1
          listgp(l1, ii) = itemp
        end if
      end do
      listgp(1, ii) = l1 - 1
      list(:, idi) = listgp(:, ii)
      phidbl(idi) = gpu_pot(ii)
      freg(1:3, idi) = gpu_a(1:3, ii)
      fdr(1:3, idi) = gpu_jerk(1:3, ii)
    end do
  end do
  call gpunb_close()

  open(11, file=trim(fout), access='stream', form='unformatted', &
       status='replace')
  write(11) rs, phidbl, freg, fdr, list, nretry
  close(11)
end program host_loop
