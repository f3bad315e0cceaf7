! ZCode: the driver of the reference (Program TermedCC, source/Main.f90) on
! the MI355X solver: read Input, build the basis, MP2 guess, then CCD or,
! with DodRPASOSEX, dRPA/SOSEX on T2aaaa (source/Main.f90:84-92); then, unless
! DoSkipTA, twist averaging (source/Main.f90:99-255, module TwistAverage).
Program TermedCC
  Use Precision
  Use Types, Only: UEGInfoType, HEGDataType
  Use IO, Only: ReadInput
  Use HEG, Only: InitHEG
  Use MP2, Only: DrvMBPT
  Use CCD, Only: DrvCCD, ShutDownSolver
  Use dRPA, Only: DrvdRPA
  Use TwistAverage, Only: DrvTwistAverage
  Implicit None
  Type (UEGInfoType) :: UEGInfo
  Type (HEGDataType) :: HEGData
  Real (Kind=pr), Allocatable :: T2aaaa(:,:,:), T2abab(:,:,:), T2abba(:,:,:)
  Integer :: IAlloc(3)

  Call ReadInput(UEGInfo)
  UEGInfo%NOcc = UEGInfo%NElectron/2
  Call InitHEG(UEGInfo, HEGData)
  Allocate(T2aaaa(UEGInfo%NOcc,UEGInfo%NOcc,UEGInfo%NOcc+1:UEGInfo%NAO), Stat=IAlloc(1))
  Allocate(T2abab(UEGInfo%NOcc,UEGInfo%NOcc,UEGInfo%NOcc+1:UEGInfo%NAO), Stat=IAlloc(2))
  Allocate(T2abba(UEGInfo%NOcc,UEGInfo%NOcc,UEGInfo%NOcc+1:UEGInfo%NAO), Stat=IAlloc(3))
  If (Any(IAlloc /= 0)) Stop 'Could not allocate in main'
  Open(7, File='Output', Status='Replace')
  Close(7)

  If (UEGInfo%DoSingleCalc) Then
    Call DrvMBPT(HEGData, UEGInfo, T2aaaa, T2abab, T2abba)
    If (.not. UEGInfo%DodRPASOSEX) Then
      Call DrvCCD(HEGData, UEGInfo, T2aaaa, T2abab, T2abba)
    Else
      Call DrvdRPA(HEGData, UEGInfo, T2aaaa)
    End If
  End If
  Call ShutDownSolver()
  If (UEGInfo%DoSkipTA) Stop 'Finishing calculation after single point calculation'
  Call DrvTwistAverage(UEGInfo, HEGData, T2aaaa, T2abab, T2abba)
End Program TermedCC
