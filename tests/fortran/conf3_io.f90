! conf3_io.f90 -- writes or reads a conf.3 snapshot with the I/O list of
! NBODY6++'s output.F:526-534 (Fortran unformatted sequential), so the
! Python reader / writer (nbody6ppgpu_amd/snapshot.py) and the driver's
! writer are checked against a real Fortran runtime.
!   conf3_io write <file> <ntot>   a file with a fixed pattern
!   conf3_io read  <file>          counts, sums and the last body
program conf3_io
  implicit none
  integer :: ntot, model, nrun, nk, j, k
  real(4), allocatable :: as(:), bodys(:), rhos(:), xns(:), xs(:,:), vs(:,:), phi(:)
  integer, allocatable :: name(:)
  character(len=512) :: mode, path, arg
  call get_command_argument(1, mode)
  call get_command_argument(2, path)
  if (trim(mode) == 'write') then
    call get_command_argument(3, arg)
    read (arg, *) ntot
    nk = 20
    model = 3
    nrun = 7
    allocate (as(nk), bodys(ntot), rhos(ntot), xns(ntot), xs(3,ntot), &
              vs(3,ntot), phi(ntot), name(ntot))
    do k = 1, nk
      as(k) = 0.5 * real(k)
    end do
    do j = 1, ntot
      bodys(j) = 1.0 / real(j)
      rhos(j) = 2.0 * real(j)
      xns(j) = real(j) + 0.25
      xs(1,j) = real(j)
      xs(2,j) = -real(j)
      xs(3,j) = 0.5 * real(j)
      vs(1,j) = 0.125 * real(j)
      vs(2,j) = 3.0
      vs(3,j) = -0.75 * real(j)
      phi(j) = -real(j)
      name(j) = 1000 + j
    end do
    open (unit=3, status='unknown', form='unformatted', file=trim(path))
    write (3) ntot, model, nrun, nk
    write (3) (as(k),k=1,nk), &
         (bodys(j),j=1,ntot), (rhos(j),j=1,ntot), (xns(j),j=1,ntot), &
         ((xs(k,j),k=1,3),j=1,ntot), ((vs(k,j),k=1,3),j=1,ntot), &
         (phi(j),j=1,ntot), (name(j),j=1,ntot)
    close (3)
  else
    open (unit=3, status='old', form='unformatted', file=trim(path))
    read (3) ntot, model, nrun, nk
    allocate (as(nk), bodys(ntot), rhos(ntot), xns(ntot), xs(3,ntot), &
              vs(3,ntot), phi(ntot), name(ntot))
    read (3) (as(k),k=1,nk), &
         (bodys(j),j=1,ntot), (rhos(j),j=1,ntot), (xns(j),j=1,ntot), &
         ((xs(k,j),k=1,3),j=1,ntot), ((vs(k,j),k=1,3),j=1,ntot), &
         (phi(j),j=1,ntot), (name(j),j=1,ntot)
    close (3)
    print '(4I12)', ntot, model, nrun, nk
    print '(ES26.17)', sum(dble(as)), sum(dble(bodys)), sum(dble(rhos)), &
         sum(dble(xns)), sum(dble(xs)), sum(dble(vs)), sum(dble(phi))
    print '(I20)', sum(int(name, 8))
    print '(ES26.17)', dble(xs(1:3,ntot)), dble(vs(1:3,ntot)), dble(bodys(ntot))
  end if
end program conf3_io
