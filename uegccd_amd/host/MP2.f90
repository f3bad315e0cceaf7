! MP2 amplitudes as the CCD guess, and the Output block of DrvMBPT
! (source/MP2.f90:11-157, formats 1000-1060); the amplitudes come from
! libuegccd (uegccd_mp2).
Module MP2
  Use, Intrinsic :: ISO_C_Binding
  Use Precision
  Use Types, Only: UEGInfoType, HEGDataType
  Use Wrappers
  Use CCD, Only: SetUpSolver, SolverContext
  Implicit None
  Private
  Public :: DrvMBPT

Contains

  Subroutine DrvMBPT(HEGData, UEGInfo, T2aaaa, T2abab, T2abba)
    Type (HEGDataType), Intent(InOut) :: HEGData
    Type (UEGInfoType), Intent(In) :: UEGInfo
    Real (Kind=pr), Intent(Out) :: T2aaaa(UEGInfo%NOcc,UEGInfo%NOcc,UEGInfo%NOcc+1:UEGInfo%NAO), &
                                   T2abab(UEGInfo%NOcc,UEGInfo%NOcc,UEGInfo%NOcc+1:UEGInfo%NAO), &
                                   T2abba(UEGInfo%NOcc,UEGInfo%NOcc,UEGInfo%NOcc+1:UEGInfo%NAO)
    Real(c_double) :: E2
    Integer :: I, A

    Write(6,*) 'Doing MBPT2...'
    If (.not. c_associated(SolverContext())) Call SetUpSolver(HEGData, UEGInfo)
    Call Check(uegccd_mp2(SolverContext(), T2aaaa, T2abab, T2abba, E2), 'uegccd_mp2')
    HEGData%ECorr = E2

    Open(7, File='Output', Position='Append')
    Write(7,1000)
    Write(7,1005) HEGData%EHF
    Write(7,1040)
    Write(7,1030)
    Write(7,1040)
    Do I = 1, UEGInfo%NOcc
      Write(7,1050) HEGData%Eigen(I)
    End Do
    Write(7,1060)
    Do A = UEGInfo%NOcc+1, UEGInfo%NAO
      Write(7,1050) HEGData%Eigen(A)
    End Do
    Write(7,1000)
    Write(7,1010) HEGData%ECorr
    Write(7,1020) HEGData%ECorr + HEGData%EHF
    Write(7,1000)
    Close(7)
    Call WriteMP2SF()

1000  Format(14x,'**************************************************')
1005  Format(14x,'* E(SCF) is ',12x,F18.12,' a.u.',2x,'*')
1010  Format(14x,'* E(2) is ',14x,F18.12,' a.u.',2x,'*')
1020  Format(14x,'* E(SCF) + E(2) is ',5x,F18.12,' a.u.',2x,'*')
1030  Format(14X,'*                RHF Eigenvalues:                *')
1040  Format(14x,'*------------------------------------------------*')
1050  Format(14x,'*',12x,ES24.16,12x,'*')
1060  Format(14x,'*',12x,24("-"),12x,'*')

    If (UEGInfo%DoOnlyMP2Grid) Stop 'DoOnlyMP2Grid flag is true, so calculation stopping here'
  End Subroutine DrvMBPT

  ! DrvMBPT's structure factor (source/MP2.f90:86-117), written to unit 81 on
  ! every call as the reference does: cells, SfG, SfSum, SfV from the library.
  Subroutine WriteMP2SF()
    Integer(c_int32_t), Allocatable, Target :: Q(:,:)
    Real(c_double), Allocatable, Target :: V(:,:)
    Integer(c_int32_t) :: N
    Real(c_double) :: E
    Integer :: K

    N = 0
    Call Check(uegccd_mp2_structure_factor(SolverContext(), N, c_null_ptr, c_null_ptr, E), &
               'uegccd_mp2_structure_factor')
    Allocate(Q(3, Max(N, 1)), V(3, Max(N, 1)))
    Call Check(uegccd_mp2_structure_factor(SolverContext(), N, c_loc(Q), c_loc(V), E), &
               'uegccd_mp2_structure_factor')
    Write(81,*) "--------"
    Write(81,*) "Starting a new structure factor"
    Write(81,*) "--------"
    Do K = 1, N
      Write(81,*) Q(1,K), Q(2,K), Q(3,K), V(1,K), V(2,K), V(3,K)
    End Do
    Flush(81)
  End Subroutine WriteMP2SF

End Module MP2
