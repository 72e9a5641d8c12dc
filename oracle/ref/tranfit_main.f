C     Builder-written main program for the reference transport fitting
C     program (devel-izem/src/tranfit/tranfit.f with its interpreter
C     ckinterp36.f): reads chem.inp, therm.dat and tran.dat in the cwd and
C     writes tranfit.out.  TEST INFRASTRUCTURE ONLY.
      PROGRAM TFMAIN
      CALL TRANFIT
      END
