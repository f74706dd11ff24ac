*     The reference host's use of the irregular-force library: implicit
*     interfaces, the per-particle SET_JP / SET_LIST calls inside an OpenMP
*     parallel loop (intgrt.F:200-205), then FIRR_VEC (intgrt.F:545).
*     Compiled on its own so the compiler sees no interface.
      SUBROUTINE LEGACY_IRR(N,LMAX,X0,X0DOT,F,FDOT,BODY,T0,LIST,TIME,
     &     NXTLEN,NXTLST,FIRR,FD,IMINR)
      INTEGER N,LMAX,NXTLEN,NXTLST(*),IMINR(*),LIST(LMAX,*)
      REAL*8 X0(3,*),X0DOT(3,*),F(3,*),FDOT(3,*),BODY(*),T0(*),TIME
      REAL*8 FIRR(3,*),FD(3,*)
      INTEGER I,RANK,NN,LM
      RANK = 0
      NN = N
      LM = LMAX
      CALL IRR_SIMD_OPEN(NN,LM,RANK)
!$omp parallel do private(I)
      DO I = 1,N
         CALL IRR_SIMD_SET_JP(I,X0(1,I),X0DOT(1,I),F(1,I),FDOT(1,I),
     &        BODY(I),T0(I))
         CALL IRR_SIMD_SET_LIST(I,LIST(1,I))
      END DO
!$omp end parallel do
      CALL IRR_SIMD_FIRR_VEC(TIME,NXTLEN,NXTLST,FIRR,FD,IMINR)
      CALL IRR_SIMD_CLOSE(RANK)
      RETURN
      END
