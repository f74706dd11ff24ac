! gpunb_iface.f90 -- ISO_C_BINDING interfaces of the MI355X regular-force
! library (include/gpunb.h).  Every procedure binds to the same symbol the
! reference's Fortran host calls through implicit interfaces (lower case,
! trailing underscore, all arguments by reference), so a caller may switch
! between "CALL GPUNB_REGF(...)" (implicit, as util_gpu.F:60 does) and the
! checked interface below without any change in the bytes passed.
module gpunb_iface
  use iso_c_binding, only: c_int, c_double, c_long_long, c_ptr
  implicit none

  ! gpunb_regcor_par / gpunb_regcor_state of include/gpunb.h: the common6.h
  ! scalars REGCOR_GPU reads, and device addresses of the body arrays.
  type, bind(C) :: gpunb_regcor_par
    real(c_double) :: time, rscale, rdens(3), rc2, rsfac, znbmin, smax, &
                      dtmin, etar, smin, bodym, tidal4
    integer(c_int) :: nnbopt, nnbmax, n, ntot, ifirst, lmax, nb_flag, m_flag
    integer(c_int) :: kz13, kz14, kz37, kz38
  end type
  type, bind(C) :: gpunb_regcor_state
    type(c_ptr) :: x0, x0dot, xn, xndot, fi, fidot, fr, frdot, d0, d0r, &
                   d1r, d2r, d3r, t0r, stepr, step, t0, timenw, rs, body, h, &
                   list
  end type

  interface
    ! src/Main/gpunb.gpu.cu:822-824
    subroutine gpunb_devinit(irank) bind(C, name="gpunb_devinit_")
      import :: c_int
      integer(c_int), intent(in) :: irank
    end subroutine
    ! src/Main/gpunb.gpu.cu:825-827 (caller intgrt.F:168, fpoly0.F:31)
    subroutine gpunb_open(nbmax, irank) bind(C, name="gpunb_open_")
      import :: c_int
      integer(c_int), intent(in) :: nbmax, irank
    end subroutine
    ! src/Main/gpunb.gpu.cu:828-830 (caller fpoly0.F:375, intgrt.F:1497)
    subroutine gpunb_close() bind(C, name="gpunb_close_")
    end subroutine
    ! src/Main/gpunb.gpu.cu:831-833 (caller intgrt.F:918, fpoly0.F:76)
    subroutine gpunb_send(nj, mj, xj, vj) bind(C, name="gpunb_send_")
      import :: c_int, c_double
      integer(c_int), intent(in) :: nj
      real(c_double), intent(in) :: mj(*), xj(3,*), vj(3,*)
    end subroutine
    ! src/Main/gpunb.gpu.cu:834-838 (caller util_gpu.F:60, fpoly0.F:125)
    subroutine gpunb_regf(ni, h2, dtr, xi, vi, acc, jrk, pot, lmax, nnbmax, &
                          list, m_flag) bind(C, name="gpunb_regf_")
      import :: c_int, c_double
      integer(c_int), intent(in) :: ni, lmax, nnbmax, m_flag
      real(c_double), intent(in) :: h2(*), dtr(*), xi(3,*), vi(3,*)
      real(c_double), intent(out) :: acc(3,*), jrk(3,*), pot(*)
      integer(c_int), intent(out) :: list(lmax,*)
    end subroutine
    ! Additive: regf plus the nearest-neighbour index (IMINR, nbint.F:29-47).
    subroutine gpunb_regfnn(ni, h2, dtr, xi, vi, acc, jrk, pot, lmax, nnbmax, &
                            list, m_flag, nnid) bind(C, name="gpunb_regfnn_")
      import :: c_int, c_double
      integer(c_int), intent(in) :: ni, lmax, nnbmax, m_flag
      real(c_double), intent(in) :: h2(*), dtr(*), xi(3,*), vi(3,*)
      real(c_double), intent(out) :: acc(3,*), jrk(3,*), pot(*)
      integer(c_int), intent(out) :: list(lmax,*), nnid(*)
    end subroutine
    ! src/Main/gpunb.gpu.cu:839-841 (caller adjust.F:315)
    subroutine gpunb_profile(irank) bind(C, name="gpunb_profile_")
      import :: c_int
      integer(c_int), intent(in) :: irank
    end subroutine
    ! src/Main/gpupot.gpu.cu:116-127 (caller energy.F:38)
    subroutine gpupot(irank, istart, ni, n, m, x, pot) bind(C, name="gpupot_")
      import :: c_int, c_double
      integer(c_int), intent(in) :: irank, istart, ni, n
      real(c_double), intent(in) :: m(*), x(3,*)
      real(c_double), intent(out) :: pot(*)
    end subroutine
    ! Additive: device-resident predictor state (irr_simd_set_jp_ in bulk).
    subroutine gpunb_jpset(n, addr, x0, v0, f, fdot, mass, t0) &
        bind(C, name="gpunb_jpset_")
      import :: c_int, c_double
      integer(c_int), intent(in) :: n, addr(*)
      real(c_double), intent(in) :: x0(3,*), v0(3,*), f(3,*), fdot(3,*)
      real(c_double), intent(in) :: mass(*), t0(*)
    end subroutine
    ! Additive: predict all stored j to time (xbpredall.f:18-26 on the GPU).
    subroutine gpunb_predict(time) bind(C, name="gpunb_predict_")
      import :: c_double
      real(c_double), intent(in) :: time
    end subroutine
    subroutine gpunb_jpget(n, x, v) bind(C, name="gpunb_jpget_")
      import :: c_int, c_double
      integer(c_int), intent(in) :: n
      real(c_double), intent(out) :: x(3,*), v(3,*)
    end subroutine
    ! Irregular force, src/Main/irr.avx.cpp:565-603 (callers intgrt.F:175,
    ! 202-204, 545, 971, 1194, 1270, 1502).  Addresses and members 1-based.
    subroutine irr_simd_open(nmax, lmax, rank) bind(C, name="irr_simd_open_")
      import :: c_int
      integer(c_int), intent(in) :: nmax, lmax, rank
    end subroutine
    subroutine irr_simd_close(rank) bind(C, name="irr_simd_close_")
      import :: c_int
      integer(c_int), intent(in) :: rank
    end subroutine
    subroutine irr_simd_profile(rank) bind(C, name="irr_simd_profile_")
      import :: c_int
      integer(c_int), intent(in) :: rank
    end subroutine
    subroutine irr_simd_set_jp(addr, pos, vel, acc2, jrk6, mass, time) &
        bind(C, name="irr_simd_set_jp_")
      import :: c_int, c_double
      integer(c_int), intent(in) :: addr
      real(c_double), intent(in) :: pos(3), vel(3), acc2(3), jrk6(3), mass, time
    end subroutine
    subroutine irr_simd_set_list(addr, nblist) bind(C, name="irr_simd_set_list_")
      import :: c_int
      integer(c_int), intent(in) :: addr, nblist(*)
    end subroutine
    subroutine irr_simd_firr_vec(ti, ni, addr, acc, jrk, nnbid) &
        bind(C, name="irr_simd_firr_vec_")
      import :: c_int, c_double
      real(c_double), intent(in) :: ti
      integer(c_int), intent(in) :: ni, addr(*)
      real(c_double), intent(out) :: acc(3,*), jrk(3,*)
      integer(c_int), intent(out) :: nnbid(*)
    end subroutine
    ! Device regular corrector for a block (replaces the REGCOR_GPU loop of
    ! util_gpu.F:152-180); d_* are device addresses.
    function gpunb_regcor_dev(par, st, ni, d_idx, d_xi, d_vi, d_freg, d_fdr, &
                              d_nnb, d_off, d_list, list_len, d_counters) &
        result(rc) bind(C, name="gpunb_regcor_dev")
      import :: c_int, c_long_long, c_ptr, gpunb_regcor_par, gpunb_regcor_state
      type(gpunb_regcor_par), intent(in) :: par
      type(gpunb_regcor_state), intent(in) :: st
      integer(c_int), value :: ni
      type(c_ptr), value :: d_idx, d_xi, d_vi, d_freg, d_fdr, d_nnb, d_off, &
                            d_list, d_counters
      integer(c_long_long), value :: list_len
      integer(c_int) :: rc
    end function
  end interface
end module gpunb_iface
