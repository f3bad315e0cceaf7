! Solver settings of the CCD iteration.  The reference's Input format stays
! untouched (an unknown keyword is an error there), so the new knobs come from
! the environment:
!   UEGCCD_MODE       reference (default: the reference's damped iteration and
!                     stopping rule, source/CCD.f90:80-99) | diis
!   UEGCCD_NDIIS      DIIS length (default 8)
!   UEGCCD_PRECISION  fp64 (default) | mixed (FP32 ladders until max|R| < 1e-5)
!   UEGCCD_TOL_E      DIIS energy tolerance (default 1e-10)
!   UEGCCD_TOL_R      DIIS residual tolerance (default 1e-8)
Module DIISCC
  Use, Intrinsic :: ISO_C_Binding
  Use Precision
  Use Wrappers
  Implicit None
  Private
  Public :: SolverOptions

Contains

  Subroutine SolverOptions(Opt)
    Type (uegccd_options), Intent(Out) :: Opt
    Character (Len=32) :: S
    Integer :: L, IOS
    Call uegccd_default_options(Opt)
    Call Get_Environment_Variable('UEGCCD_MODE', S, L)
    If (L > 0 .and. (S(1:4) == 'diis' .or. S(1:4) == 'DIIS')) Opt%mode = UEGCCD_MODE_DIIS
    Call Get_Environment_Variable('UEGCCD_NDIIS', S, L)
    If (L > 0) Then
      Read(S,*,IOStat=IOS) Opt%ndiis
    End If
    Call Get_Environment_Variable('UEGCCD_PRECISION', S, L)
    If (L > 0 .and. S(1:5) == 'mixed') Opt%precision = UEGCCD_PREC_MIXED
    Call Get_Environment_Variable('UEGCCD_TOL_E', S, L)
    If (L > 0) Then
      Read(S,*,IOStat=IOS) Opt%tol_e
    End If
    Call Get_Environment_Variable('UEGCCD_TOL_R', S, L)
    If (L > 0) Then
      Read(S,*,IOStat=IOS) Opt%tol_r
    End If
  End Subroutine SolverOptions

End Module DIISCC
