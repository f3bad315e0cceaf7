! Working precision of the host (the reference uses real(8) throughout,
! source/Precision.f90:5-6).
Module Precision
  Use, Intrinsic :: ISO_Fortran_Env, Only: real64
  Implicit None
  Integer, Parameter :: pr = real64
  Real (Kind=pr), Parameter :: Zero = 0.0_pr, One = 1.0_pr, F12 = 0.5_pr
End Module Precision
