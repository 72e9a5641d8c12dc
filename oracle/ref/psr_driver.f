C     psr_driver.f -- builder-written harness around the reference CHEMKIN-II
C     and DVODE sources (devel-izem/src/chemkin/*.f of saras13015/Phd), which
C     are compiled where they lie.  TEST INFRASTRUCTURE ONLY.
C
C     Entry points (called through ctypes, all arguments by reference):
C       REFINIT            : INITCHEMKIN(.FALSE.) -- reads chem.bin in the cwd
C       REFPSR  (T,RHO,P,Y,DT,ISTATE)      : PSR_CHEMKIN, COSINESS 7-argument
C                            call form (cos/cosiness/module_reaction.f90:294),
C                            reference tolerances RTOL 1e-6 / ATOL 1e-10
C       REFPSRT (T,RHO,Y,DT,RTOLI,ATOLI,ISTATE,NST,NFE,NJE)
C                          : the loop of REACTION_V (reaction_v.f:14-42)
C                            with caller tolerances; returns DVODE counters
C       REFWDOT (RHO,T,Y,WDOT), REFHMS (T,H), REFCVB (T,Y,CVB),
C       REFRHS  (RHO,Z,ZP) : CKWYR, CKHMS, CKCVBS, FUN
C       REFKK   (KK,NEQO)  : species / equation counts
      SUBROUTINE REFINIT
      CALL INITCHEMKIN (.FALSE.)
      END

      SUBROUTINE REFKK (KKO, NEQO)
      IMPLICIT NONE
      INCLUDE 'chemkin.common'
      INTEGER KKO, NEQO
      KKO = KK
      NEQO = NEQ
      END

      SUBROUTINE REFPSR (T, RHO, P, Y, DT, ISTATE)
      IMPLICIT NONE
      DOUBLE PRECISION T, RHO, P, Y(*), DT
      INTEGER ISTATE
      CALL PSR_CHEMKIN (T, RHO, P, Y, DT, 0.0D0, ISTATE)
      END

      SUBROUTINE REFPSRT (T, RHO, Y, DT, RTOLI, ATOLI, ISTATE,
     *                    NST, NFE, NJE)
      IMPLICIT NONE
      INCLUDE 'chemkin.common'
      DOUBLE PRECISION T, RHO, Y(*), DT, RTOLI, ATOLI, TI, TF
      INTEGER ISTATE, NST, NFE, NJE, IB, K, MF
      EXTERNAL FUN, JAC
      RHO_ = RHO
      OMS  = 1.0D0
      Z(1) = T
      DO K = 2, KK + 1
         Z(K) = Y(K-1)
      ENDDO
      TI = 0.0D0
      TF = DT
      MF = 22
      IB = 0
      ISTATE = 1
      NST = 0
      NFE = 0
      NJE = 0
      DO WHILE ( (ISTATE .NE. 2) .AND. (IB .LT. 20) )
         ISTATE = 1
         IB = IB + 1
         CALL DVODE ( FUN, NEQ, Z, TI, TF, ITOL, RTOLI, ATOLI, ITASK,
     *                ISTATE, IOPT, RWORK(NVODE), LRW, IWORK(IVODE),
     *                LIW, JAC, MF, RWORK, IWORK )
         NST = NST + IWORK(IVODE + 10)
         NFE = NFE + IWORK(IVODE + 11)
         NJE = NJE + IWORK(IVODE + 12)
      END DO
      T = Z(1)
      DO K = 1, KK
         Y(K) = Z(K+1)
      ENDDO
      END

      SUBROUTINE REFWDOT (RHO, T, Y, WDOT)
      IMPLICIT NONE
      INCLUDE 'chemkin.common'
      DOUBLE PRECISION RHO, T, Y(*), WDOT(*)
      CALL CKWYR (RHO, T, Y, IWORK, RWORK, WDOT)
      END

      SUBROUTINE REFHMS (T, H)
      IMPLICIT NONE
      INCLUDE 'chemkin.common'
      DOUBLE PRECISION T, H(*)
      CALL CKHMS (T, IWORK, RWORK, H)
      END

      SUBROUTINE REFCVB (T, Y, CVB)
      IMPLICIT NONE
      INCLUDE 'chemkin.common'
      DOUBLE PRECISION T, Y(*), CVB
      CALL CKCVBS (T, Y, IWORK, RWORK, CVB)
      END

      SUBROUTINE REFRHS (RHO, ZI, ZPO)
      IMPLICIT NONE
      INCLUDE 'chemkin.common'
      DOUBLE PRECISION RHO, ZI(*), ZPO(*), TT
      RHO_ = RHO
      OMS  = 1.0D0
      TT = 0.0D0
      CALL FUN (NEQ, TT, ZI, ZPO, RWORK, IWORK)
      END
