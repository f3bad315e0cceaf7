! The dRPA driver of the host, DrvdRPA (source/dRPA.f90:23-143): the ring-only
! CCD on the spin-summed T2 runs in libuegccd (uegccd_drpa); this routine binds
! the context, calls it and writes the reference's Output block (formats of
! source/dRPA.f90:122-139).  HEGData%ECorr ends as the reference leaves it:
! Ec(SOSEX) after a converged run, Ec(dRPA) otherwise.
Module dRPA
  Use, Intrinsic :: ISO_C_Binding
  Use Precision
  Use Types, Only: UEGInfoType, HEGDataType
  Use Wrappers
  Use CCD, Only: SetUpSolver, SolverContext
  Implicit None
  Private
  Public :: DrvdRPA

Contains

  Subroutine DrvdRPA(HEGData, UEGInfo, T2)
    Type (HEGDataType), Intent(InOut) :: HEGData
    Type (UEGInfoType), Intent(In) :: UEGInfo
    Real (Kind=pr), Intent(InOut) :: T2(UEGInfo%NOcc,UEGInfo%NOcc,UEGInfo%NOcc+1:UEGInfo%NAO)
    Type (uegccd_result) :: Res
    Integer(c_int) :: St
    Real(c_double) :: ESosex
    Real(c_double), Allocatable :: E(:), DT(:)
    Integer :: K

    Write(6,*) 'Doing RPA...'
    If (.not. c_associated(SolverContext())) Call SetUpSolver(HEGData, UEGInfo)
    St = uegccd_drpa(SolverContext(), T2, Res, ESosex)
    If (St /= UEGCCD_OK .and. St /= UEGCCD_ERR_NOT_CONVERGED) Call Check(St, 'uegccd_drpa')
    HEGData%ECorr = Res%ecorr
    Allocate(E(Max(Res%niter,1)), DT(Max(Res%niter,1)))
    Call Check(uegccd_trace(SolverContext(), E, DT, Res%niter), 'uegccd_trace')

    Open(7, File='Output', Position='Append')
    Write(7,1010)
    Write(7,1020)
    If (UEGInfo%DoRing)   Write(7,1100)
    If (UEGInfo%DoMosaic) Write(7,1130)
    Write(7,1020)
    Write(7,1030)
    Do K = 1, Res%niter
      Write(7,1040) E(K), K, DT(K)
      If (K > 1) Write(6,1040) E(K), K, DT(K)
    End Do
    Write(7,1020)
    If (Res%failed /= 0) Then
      Write(7,2000)
    Else
      Write(7,1050) Res%niter
      Write(7,1060) HEGData%ECorr
      Write(7,1070) HEGData%ECorr + HEGData%EHF
      HEGData%ECorr = ESosex
      Write(7,1065) HEGData%ECorr
      Write(7,1075) HEGData%ECorr + HEGData%EHF
    End If
    Write(7,1090) Res%maxres
    Write(7,1000)
    Close(7)

1000  Format(14x,'**************************************************')
1010  Format(14X,'*              dRPA summary follows              *')
1020  Format(14x,'*------------------------------------------------*')
1030  Format(14X,'* dRPA Energy      Iteration    Biggest change   *')
1040  Format(14X,'* ',F15.10,2x,I5,7X,F14.10,4x,'*')
1050  Format(14x,'* dRPA has converged in ',I3,' iterations',11x,'*')
1060  Format(14x,'* Ec(dRPA) is ',9x,F18.12,' a.u.   *')
1065  Format(14x,'* Ec(SOSEX) is ',8x,F18.12,' a.u.   *')
1070  Format(14x,'* Final dRPA Energy is ',F18.12,' a.u.   *')
1075  Format(14x,'* Final SOSEX Energy is',F18.12,' a.u.   *')
1090  Format(14x,'* Max CCD Residual is  ',4x,ES18.12,'    *')
1100  Format(14x,'* Including ring diagrams...                     *')
1130  Format(14x,'* Including mosaic diagrams...                   *')
2000  Format(14x,'* Final dRPA Energy is   DID NOT CONVERGE        *')
  End Subroutine DrvdRPA

End Module dRPA
