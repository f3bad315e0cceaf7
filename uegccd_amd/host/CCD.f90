! The CCD driver of the host, DrvCCD (source/CCD.f90:18-183): the iteration
! itself runs in libuegccd; this routine binds the context, calls the solver
! and writes the reference's Output block (formats of source/CCD.f90:127-140).
Module CCD
  Use, Intrinsic :: ISO_C_Binding
  Use Precision
  Use Types, Only: UEGInfoType, HEGDataType
  Use Wrappers
  Use DIISCC, Only: SolverOptions
  Implicit None
  Private
  Public :: SetUpSolver, ShutDownSolver, DrvCCD, SolverContext

  Type(c_ptr), Save :: Ctx = c_null_ptr

Contains

  Function SolverContext() Result(C)
    Type(c_ptr) :: C
    C = Ctx
  End Function SolverContext

  ! uegccd_ctx_create for the current basis (index maps + device buffers).
  Subroutine SetUpSolver(HEGData, UEGInfo)
    Type (HEGDataType), Intent(In) :: HEGData
    Type (UEGInfoType), Intent(In) :: UEGInfo
    Type (uegccd_system) :: Sys
    Type (uegccd_options) :: Opt
    Integer(c_int32_t), Allocatable, Target :: NVecC(:)
    Real(c_double), Allocatable, Target :: Eig(:)

    Call ShutDownSolver()
    Allocate(NVecC(3*UEGInfo%NAO), Eig(UEGInfo%NAO))
    NVecC = Reshape(HEGData%nVec, [3*UEGInfo%NAO])
    Eig = HEGData%Eigen
    Sys%nocc = UEGInfo%NOcc
    Sys%nao = UEGInfo%NAO
    Sys%nvec = c_loc(NVecC)
    Sys%eigen = c_loc(Eig)
    Sys%L = HEGData%L
    Sys%madelung = HEGData%Madelung
    Sys%rs = HEGData%rS
    Sys%gapii = UEGInfo%GapII
    Sys%coren = UEGInfo%CoreN
    Sys%safe_eri = Merge(1, 0, UEGInfo%SafeERI)
    Sys%irange = UEGInfo%IRange
    Sys%do_ring = Merge(1, 0, UEGInfo%DoRing)
    Sys%do_xring = Merge(1, 0, UEGInfo%DoXRing)
    Sys%do_ladder = Merge(1, 0, UEGInfo%DoLadder)
    Sys%do_mosaic = Merge(1, 0, UEGInfo%DoMosaic)
    Sys%screening = HEGData%ScreeningDistance
    Call SolverOptions(Opt)
    Call Check(uegccd_ctx_create(Sys, Opt, Ctx), 'uegccd_ctx_create')
  End Subroutine SetUpSolver

  Subroutine ShutDownSolver()
    If (c_associated(Ctx)) Call uegccd_ctx_destroy(Ctx)
    Ctx = c_null_ptr
  End Subroutine ShutDownSolver

  Subroutine DrvCCD(HEGData, UEGInfo, T2aaaa, T2abab, T2abba)
    Type (HEGDataType), Intent(InOut) :: HEGData
    Type (UEGInfoType), Intent(In) :: UEGInfo
    Real (Kind=pr), Intent(InOut) :: T2aaaa(UEGInfo%NOcc,UEGInfo%NOcc,UEGInfo%NOcc+1:UEGInfo%NAO), &
                                     T2abab(UEGInfo%NOcc,UEGInfo%NOcc,UEGInfo%NOcc+1:UEGInfo%NAO), &
                                     T2abba(UEGInfo%NOcc,UEGInfo%NOcc,UEGInfo%NOcc+1:UEGInfo%NAO)
    Type (uegccd_result) :: Res
    Integer(c_int) :: St
    Real(c_double), Allocatable :: E(:), DT(:)
    Integer :: K

    Write(6,*) 'Doing CCD...'
    If (.not. c_associated(Ctx)) Call SetUpSolver(HEGData, UEGInfo)
    If (UEGInfo%DoSFCalcMP2) Call WriteSF(T2abab, T2abba)
    St = uegccd_solve(Ctx, T2aaaa, T2abab, T2abba, Res)
    If (St /= UEGCCD_OK .and. St /= UEGCCD_ERR_NOT_CONVERGED) Call Check(St, 'uegccd_solve')
    HEGData%ECorr = Res%ecorr
    If (UEGInfo%DoSFCalcCCD) Call WriteSF(T2abab, T2abba)
    Allocate(E(Max(Res%niter,1)), DT(Max(Res%niter,1)))
    Call Check(uegccd_trace(Ctx, E, DT, Res%niter), 'uegccd_trace')

    Open(7, File='Output', Position='Append')
    Write(7,1010)
    Write(7,1020)
    If (UEGInfo%DoRing)   Write(7,1100)
    If (UEGInfo%DoXRing)  Write(7,1110)
    If (UEGInfo%DoLadder) Write(7,1120)
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
    End If
    Write(7,1090) Res%maxres
    Write(7,1000)
    Close(7)

1000  Format(14x,'**************************************************')
1010  Format(14X,'*               CCD summary follows              *')
1020  Format(14x,'*------------------------------------------------*')
1030  Format(14X,'*  CCD Energy      Iteration    Biggest change   *')
1040  Format(14X,'* ',F15.10,2x,I5,7X,F14.10,4x,'*')
1050  Format(14x,'* CCD has converged in ',I3,' iterations',12x,'*')
1060  Format(14x,'* Ec(CCD) is ',10x,F18.12,' a.u.   *')
1070  Format(14x,'* Final CCD Energy is  ',F18.12,' a.u.   *')
1090  Format(14x,'* Max CCD Residual is  ',4x,ES18.12,'    *')
1100  Format(14x,'* Including ring diagrams...                     *')
1110  Format(14x,'* Including crossed ring diagrams...             *')
1120  Format(14x,'* Including ladder diagrams...                   *')
1130  Format(14x,'* Including mosaic diagrams...                   *')
2000  Format(14x,'* Final CCD Energy is   DID NOT CONVERGE         *')
  End Subroutine DrvCCD

  ! CCEnergyAndSF (source/CCD.f90:240-351): the structure factor of the
  ! amplitudes from the library, written to unit 80 in the reference's order
  Subroutine WriteSF(T2abab, T2abba)
    Real (Kind=pr), Intent(In) :: T2abab(*), T2abba(*)
    Integer(c_int32_t), Allocatable, Target :: Q(:,:)
    Real(c_double), Allocatable, Target :: V(:,:)
    Integer(c_int32_t) :: N
    Real(c_double) :: E
    Integer :: K

    N = 0
    Call Check(uegccd_structure_factor(Ctx, T2abab, T2abba, N, c_null_ptr, c_null_ptr, E), &
               'uegccd_structure_factor')
    Allocate(Q(3, Max(N, 1)), V(5, Max(N, 1)))
    Call Check(uegccd_structure_factor(Ctx, T2abab, T2abba, N, c_loc(Q), c_loc(V), E), &
               'uegccd_structure_factor')
    Write(80,*) "--------"
    Write(80,*) "Starting a new structure factor"
    Write(80,*) "--------"
    Do K = 1, N
      Write(80,*) Q(1,K), Q(2,K), Q(3,K), V(1,K), V(2,K), V(3,K), V(4,K), V(5,K)
    End Do
    Flush(80)
  End Subroutine WriteSF

End Module CCD
