C     Builder-written main program for the reference CHEMKIN-II interpreter
C     (devel-izem/src/chemkin/ckinterp39.f): reads chem.inp and therm.dat in
C     the cwd, writes chem.bin and chem.out.  TEST INFRASTRUCTURE ONLY.
      PROGRAM CKIMAIN
      CALL CKINTP39
      END
