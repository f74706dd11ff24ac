*     The reference host's spelling of the regular-force call: an implicit
*     interface, no module, as util_gpu.F:60 writes it.  Compiled on its
*     own so the compiler sees no interface for GPUNB_REGF.
      SUBROUTINE LEGACY_REGF(NI,H2,DTR,XI,VI,ACC,JRK,POT,LMAX,NNBMAX,
     &     LIST,MFLAG)
      INTEGER NI,LMAX,NNBMAX,MFLAG
      REAL*8 H2(*),DTR(*),XI(3,*),VI(3,*),ACC(3,*),JRK(3,*),POT(*)
      INTEGER LIST(LMAX,*)
      CALL GPUNB_REGF(NI,H2,DTR,XI,VI,ACC,JRK,POT,LMAX,NNBMAX,
     &     LIST,MFLAG)
      RETURN
      END
