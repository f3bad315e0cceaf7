! Twist averaging (source/Main.f90:99-255) on libuegccd: for every twist the
! plane-wave basis, E_HF, the MP2 guess, E(2) and the MP2 connectivity on the
! GPU; with DoCalcTACCD every twist's CCD, iterated together by the library's
! request server (dRPA/SOSEX twist by twist); the reference's statistics in
! fort.59 and the per-twist connectivity in fort.58; then the special twist
! (the first minimum of sum_k ((C_t(k) - <C>(k)) / k^2)^2) and the MP2 and
! CCD there with the twist-averaged eigenvalues.  Records carry the bytes of
! the reference's list-directed output (gfortran's: each item a blank and a
! 25-character field, 17 significant digits; integers 12 characters wide).
!
! The reference draws the twists with rand() seeded 86456
! (source/HEG.f90:535-549); here they are read from the file named by the
! environment variable UEGCCD_TWISTS (one twist per line) or, without it,
! drawn in GenerateTwist's form (a uniform number minus 0.5 per component)
! from the Fortran generator.
Module TwistAverage
  Use, Intrinsic :: ISO_C_Binding
  Use Precision
  Use Types, Only: UEGInfoType, HEGDataType
  Use Wrappers
  Use HEG, Only: InitHEG
  Use MP2, Only: DrvMBPT
  Use CCD, Only: DrvCCD, SetUpSolver, ShutDownSolver, SolverContext
  Use dRPA, Only: DrvdRPA
  Use DIISCC, Only: SolverOptions
  Implicit None
  Private
  Public :: DrvTwistAverage

  Integer, Parameter :: ConnectivityLimit = 400  ! source/MP2.f90:26

Contains

  Subroutine TwistList(UEGInfo, Twists, N)
    Type (UEGInfoType), Intent(In) :: UEGInfo
    Real (Kind=pr), Allocatable, Intent(Out) :: Twists(:,:)
    Integer, Intent(Out) :: N
    Character (Len=512) :: Path
    Integer :: L, IOS, K, U
    Real (Kind=pr) :: T(3)
    Real :: R(3)
    Integer, Allocatable :: Seed(:)

    Call Get_Environment_Variable('UEGCCD_TWISTS', Path, L)
    If (L > 0) Then
      Open(NewUnit=U, File=Path(1:L), Status='Old', Action='Read', IOStat=IOS)
      If (IOS /= 0) Stop 'UEGCCD_TWISTS: cannot open the twist file'
      N = 0
      Do
        Read(U, *, IOStat=IOS) T
        If (IOS /= 0) Exit
        N = N + 1
      End Do
      Rewind(U)
      Allocate(Twists(3, N))
      Do K = 1, N
        Read(U, *) Twists(:, K)
      End Do
      Close(U)
    Else
      N = UEGInfo%TACalcN
      Allocate(Twists(3, N))
      Call Random_Seed(Size=L)
      Allocate(Seed(L))
      Seed = 86456
      Call Random_Seed(Put=Seed)
      Do K = 1, N
        Call Random_Number(R)
        Twists(:, K) = Real(R, pr) - 0.5_pr
      End Do
    End If
    If (N < 2) Stop 'Twist averaging needs at least two twists'
  End Subroutine TwistList

  ! Every twist's CCD through the request server: each request builds its own
  ! basis at its twist; the results come back by id.
  Subroutine ServedCCD(UEGInfo, Twists, N, Ecorr)
    Type (UEGInfoType), Intent(In) :: UEGInfo
    Real (Kind=pr), Intent(In) :: Twists(3, N)
    Integer, Intent(In) :: N
    Real (Kind=pr), Intent(Out) :: Ecorr(N)
    Type (uegccd_server_options) :: SOpt
    Type (uegccd_request) :: Req
    Type (uegccd_request_result) :: Res
    Type (c_ptr) :: Srv
    Integer(c_int64_t), Allocatable :: Id(:)
    Integer :: K

    SOpt%device = 0
    SOpt%hbm_budget_bytes = 0
    SOpt%max_active = Min(64, N)
    SOpt%iters_per_step = 4
    SOpt%planner_threads = 0
    Call Check(uegccd_server_create(SOpt, Srv), 'uegccd_server_create')
    Req%heg%nel = UEGInfo%NElectron
    Req%heg%maxkpoint = UEGInfo%MaxKPoint
    Req%heg%rs = UEGInfo%rSMin
    Req%heg%fcut = UEGInfo%fcut
    Req%heg%corefcut = UEGInfo%corefcut
    Req%heg%madelungfactor = UEGInfo%MadelungFactor
    Req%heg%gap = UEGInfo%Gap
    Req%heg%exchangefactor = UEGInfo%ExchangeFactor
    Req%heg%cutoff_twisted = 0
    Req%gapii = UEGInfo%GapII
    Req%coren = UEGInfo%CoreN
    Req%safe_eri = Merge(1, 0, UEGInfo%SafeERI)
    Req%irange = UEGInfo%IRange
    Req%do_ring = Merge(1, 0, UEGInfo%DoRing)
    Req%do_xring = Merge(1, 0, UEGInfo%DoXRing)
    Req%do_ladder = Merge(1, 0, UEGInfo%DoLadder)
    Req%do_mosaic = Merge(1, 0, UEGInfo%DoMosaic)
    Call SolverOptions(Req%opt)
    Allocate(Id(N))
    Do K = 1, N
      Req%heg%twist = Twists(:, K)
      Call Check(uegccd_server_submit(Srv, Req, Id(K)), 'uegccd_server_submit')
    End Do
    Call Check(uegccd_server_drain(Srv), 'uegccd_server_drain')
    Do K = 1, N
      Call Check(uegccd_server_poll(Srv, Id(K), Res), 'uegccd_server_poll')
      If (Res%status /= UEGCCD_OK .and. Res%status /= UEGCCD_ERR_NOT_CONVERGED) &
          Call Check(Res%status, 'served CCD of a twist')
      Ecorr(K) = Res%ecorr
    End Do
    Call uegccd_server_destroy(Srv)
  End Subroutine ServedCCD

  ! One real item of list-directed output: a blank, then 25 characters with
  ! 17 significant digits, F form for zero and 0.1 <= |x| < 1e17 (after the
  ! rounding to 17 digits), else one digit before the point and a
  ! three-digit exponent.
  Function Item(X) Result(S)
    Real (Kind=pr), Intent(In) :: X
    Character (Len=26) :: S
    Integer :: Ex
    Write(S, '(1X,1PE25.16E3)') X
    Read(S(23:26), *) Ex
    If (X == 0 .or. (Ex >= -1 .and. Ex <= 16)) Write(S, '(1X,G25.17E3)') X
  End Function Item

  Subroutine WriteRecord(U, X, Head, Tail)
    Integer, Intent(In) :: U
    Real (Kind=pr), Intent(In) :: X(:)
    Character (Len=*), Intent(In), Optional :: Head, Tail
    Character (Len=:), Allocatable :: Line
    Integer :: K
    Line = ''
    If (Present(Head)) Line = ' ' // Head
    Do K = 1, Size(X)
      Line = Line // Item(X(K))
    End Do
    If (Present(Tail)) Line = Line // ' ' // Tail
    Write(U, '(A)') Line
  End Subroutine WriteRecord

  Subroutine DrvTwistAverage(UEGInfo, HEGData, T2aaaa, T2abab, T2abba)
    Type (UEGInfoType), Intent(InOut) :: UEGInfo
    Type (HEGDataType), Intent(InOut) :: HEGData
    Real (Kind=pr), Intent(InOut) :: T2aaaa(UEGInfo%NOcc,UEGInfo%NOcc,UEGInfo%NOcc+1:UEGInfo%NAO), &
                                     T2abab(UEGInfo%NOcc,UEGInfo%NOcc,UEGInfo%NOcc+1:UEGInfo%NAO), &
                                     T2abba(UEGInfo%NOcc,UEGInfo%NOcc,UEGInfo%NOcc+1:UEGInfo%NAO)
    Real (Kind=pr), Allocatable :: Twists(:,:), Eig(:,:), HF(:), MP2(:), CCDE(:), Served(:)
    Integer(c_int32_t), Allocatable :: Conn(:,:)
    Integer(c_int64_t) :: ConnSum(ConnectivityLimit)
    Real (Kind=pr) :: NT(3), NT2(3), NTErr(3), HfA, Mp2A, CcdA, Hf2, Mp22, Ccd2, HfErr, Mp2Err, CcdErr
    Real (Kind=pr) :: ConnAvg(ConnectivityLimit), Diff(ConnectivityLimit), D2, MinD2
    Real (Kind=pr), Allocatable :: EigSum(:), EigAvg(:)
    Integer :: N, K, M, NAO, ConnMax, Special
    Integer(c_int) :: St

    Call TwistList(UEGInfo, Twists, N)
    NAO = UEGInfo%NAO
    Allocate(Eig(NAO, N), HF(N), MP2(N), CCDE(N), Conn(ConnectivityLimit, N), Served(N))
    Allocate(EigSum(NAO), EigAvg(NAO))
    If (UEGInfo%DoCalcTACCD .and. .not. UEGInfo%DodRPASOSEX) Call ServedCCD(UEGInfo, Twists, N, Served)
    ConnMax = 0
    Do K = 1, N
      Write(60, '(1X,I11,1X,A,1X,I11)') K, "of", N
      Call InitHEG(UEGInfo, HEGData, Twists(:, K))
      If (UEGInfo%NAO /= NAO) Stop 'The basis size changed with the twist'
      If (K == 1) Write(60, '(1X,A,1X,I11)') "EIGEN", NAO
      Call ShutDownSolver()
      Call DrvMBPT(HEGData, UEGInfo, T2aaaa, T2abab, T2abba)
      St = uegccd_connectivity(SolverContext(), Conn(:, K), Int(ConnectivityLimit, c_int32_t))
      If (St == -1) Then
        Print *, "Connectivity greater than current ConnectivityLimit"
        Print *, "Increase ConnectivityLimit"
        Stop "Program stopped in MP2.f90"
      End If
      Call Check(St, 'uegccd_connectivity')
      Eig(:, K) = HEGData%Eigen
      HF(K) = HEGData%EHF
      MP2(K) = HEGData%ECorr
      CCDE(K) = HEGData%ECorr
      If (UEGInfo%DoCalcTACCD) Then
        If (UEGInfo%DodRPASOSEX) Then
          Call DrvdRPA(HEGData, UEGInfo, T2aaaa)
          CCDE(K) = HEGData%ECorr
        Else
          CCDE(K) = Served(K)
        End If
      End If
      Call WriteRecord(59, [Twists(:, K), HF(K), MP2(K), CCDE(K)])
      If (K == 1) Then
        Do M = ConnectivityLimit, 1, -1
          If (Conn(M, 1) > 0) Then
            ConnMax = M
            Exit
          End If
        End Do
      End If
      If (ConnMax > 0) Write(58, '(1X,*(I11,:,1X))') Conn(1:ConnMax, K)
    End Do
    ! sums in twist order, as the reference accumulates them
    NT = 0; NT2 = 0; HfA = 0; Mp2A = 0; CcdA = 0; Hf2 = 0; Mp22 = 0; Ccd2 = 0
    ConnSum = 0
    EigSum = 0
    Do K = 1, N
      NT = NT + Twists(:, K)
      HfA = HfA + HF(K)
      Mp2A = Mp2A + MP2(K)
      CcdA = CcdA + CCDE(K)
      ConnSum = ConnSum + Conn(:, K)
      EigSum = EigSum + Eig(:, K)
      NT2 = NT2 + Twists(:, K)**2.0_pr
      Hf2 = Hf2 + HF(K)**2.0_pr
      Mp22 = Mp22 + MP2(K)**2.0_pr
      Ccd2 = Ccd2 + CCDE(K)**2.0_pr
    End Do
    NT = NT / N; HfA = HfA / N; Mp2A = Mp2A / N; CcdA = CcdA / N
    ConnAvg = Real(ConnSum, pr) / N
    EigAvg = EigSum / N
    NT2 = NT2 / N; Hf2 = Hf2 / N; Mp22 = Mp22 / N; Ccd2 = Ccd2 / N
    NTErr = ((NT2 - NT**2.0_pr) / (N - 1))**0.5_pr
    HfErr = ((Hf2 - HfA**2.0_pr) / (N - 1))**0.5_pr
    Mp2Err = ((Mp22 - Mp2A**2.0_pr) / (N - 1))**0.5_pr
    CcdErr = ((Ccd2 - CcdA**2.0_pr) / (N - 1))**0.5_pr
    Call WriteRecord(59, [NT, HfA, Mp2A, CcdA], Tail="averages")
    Call WriteRecord(59, [NT2, Hf2, Mp22, Ccd2], Tail="squared averages")
    Call WriteRecord(59, [NTErr, HfErr, Mp2Err, CcdErr], Tail="errors")
    Call WriteRecord(59, EigAvg, Tail="eigenvalue averages")
    M = 0
    Do K = ConnectivityLimit, 1, -1
      If (ConnSum(K) > 0) Then
        M = K
        Exit
      End If
    End Do
    Call WriteRecord(59, ConnAvg(1:M), Tail="conn")
    ! the twist whose connectivity is closest to the average
    Special = 1
    MinD2 = 0
    Do K = 1, N
      Do M = 1, ConnectivityLimit
        Diff(M) = (Conn(M, K) - ConnAvg(M)) / M**2.0_pr
      End Do
      D2 = Dot_Product(Diff, Diff)
      If (K == 1 .or. D2 < MinD2) Then
        MinD2 = D2
        Special = K
      End If
      Write(59, '(1X,I11,A)') K, Item(D2)
    End Do
    Call WriteRecord(59, Twists(:, Special), Head="special twist angle found")
    Call WriteRecord(59, [MinD2], Head="lowest connectivity diff squared")
    Flush(59)
    Flush(58)
    ! MP2 and CCD at the special twist with the twist-averaged eigenvalues
    Call InitHEG(UEGInfo, HEGData, Twists(:, Special))
    HEGData%Eigen = EigAvg
    Call ShutDownSolver()
    Call DrvMBPT(HEGData, UEGInfo, T2aaaa, T2abab, T2abba)
    If (.not. UEGInfo%DodRPASOSEX) Then
      Call DrvCCD(HEGData, UEGInfo, T2aaaa, T2abab, T2abba)
    Else
      Call DrvdRPA(HEGData, UEGInfo, T2aaaa)
    End If
    Call ShutDownSolver()
  End Subroutine DrvTwistAverage

End Module TwistAverage
