! Plane-wave basis of the single-point calculation, built on the GPU by
! libuegccd (uegccd_heg_build_device: init_HEG + SetupBasis + change_rs of
! source/HEG.f90).
Module HEG
  Use, Intrinsic :: ISO_C_Binding
  Use Precision
  Use Types, Only: UEGInfoType, HEGDataType
  Use Wrappers
  Implicit None
  Private
  Public :: InitHEG

Contains

  ! Twist, when given, replaces the Input's NTwist: the basis of
  ! ChangeTwist + SetupBasis at the current r_s (source/HEG.f90:551-578), which
  ! prints one exchange structure factor; without it, init_HEG + change_rs,
  ! which print two (at r_s = 1 and at r_s).
  Subroutine InitHEG(UEGInfo, HEGData, Twist)
    Type (UEGInfoType), Intent(InOut) :: UEGInfo
    Type (HEGDataType), Intent(InOut) :: HEGData
    Real (Kind=pr), Intent(In), Optional :: Twist(3)
    Type (uegccd_heg_params) :: P
    Integer(c_int32_t) :: NAO
    Integer(c_int32_t), Allocatable, Target :: NVecC(:)
    Real(c_double), Allocatable, Target :: Eig(:)
    Real(c_double), Target :: Scal(8)

    P%nel = UEGInfo%NElectron
    P%maxkpoint = UEGInfo%MaxKPoint
    P%rs = UEGInfo%rSMin
    P%twist = UEGInfo%NTwist
    If (Present(Twist)) P%twist = Twist
    P%fcut = UEGInfo%fcut
    P%corefcut = UEGInfo%corefcut
    P%madelungfactor = UEGInfo%MadelungFactor
    P%gap = UEGInfo%Gap
    P%exchangefactor = UEGInfo%ExchangeFactor
    P%cutoff_twisted = 0
    Call Check(uegccd_heg_build_device(P, 0_c_int32_t, NAO, c_null_ptr, c_null_ptr, c_null_ptr), &
               'uegccd_heg_build_device')
    Allocate(NVecC(3*NAO), Eig(NAO))
    Call Check(uegccd_heg_build_device(P, 0_c_int32_t, NAO, c_loc(NVecC), c_loc(Eig), c_loc(Scal)), &
               'uegccd_heg_build_device')
    UEGInfo%NAO = NAO
    HEGData%Nel = UEGInfo%NElectron
    HEGData%rS = UEGInfo%rSMin
    HEGData%L = Scal(1)
    HEGData%Omega = Scal(2)
    HEGData%Madelung = Scal(3)
    HEGData%EHF = Scal(4)
    HEGData%ScreeningDistance = Scal(5)
    HEGData%DenomFactor = Scal(6)
    HEGData%TolMax = Scal(7)
    HEGData%FailRatio = Scal(8)
    If (Allocated(HEGData%nVec)) Deallocate(HEGData%nVec)
    If (Allocated(HEGData%Eigen)) Deallocate(HEGData%Eigen)
    Allocate(HEGData%nVec(NAO,3), HEGData%Eigen(NAO))
    HEGData%nVec = Reshape(NVecC, [Int(NAO), 3])
    HEGData%Eigen = Eig
    ! print_HF's exchange structure factor: once for the set-up at r_s = 1
    ! and once after the change to r_s (source/HEG.f90:319,531,613)
    If (.not. Present(Twist)) Call WriteExSF(P, 1.0_c_double)
    Call WriteExSF(P, UEGInfo%rSMin)
  End Subroutine InitHEG

  ! print_ex_SF (source/HEG.f90:617-674) to unit 82, from the library.
  Subroutine WriteExSF(P0, RS)
    Type (uegccd_heg_params), Intent(In) :: P0
    Real(c_double), Intent(In) :: RS
    Type (uegccd_heg_params) :: P
    Integer(c_int32_t), Allocatable, Target :: Q(:,:), Cnt(:)
    Real(c_double), Allocatable, Target :: V(:)
    Integer(c_int32_t) :: N
    Integer :: K

    P = P0
    P%rs = RS
    N = 0
    Call Check(uegccd_exchange_structure_factor(P, N, c_null_ptr, c_null_ptr, c_null_ptr), &
               'uegccd_exchange_structure_factor')
    Allocate(Q(3, Max(N, 1)), Cnt(Max(N, 1)), V(Max(N, 1)))
    Call Check(uegccd_exchange_structure_factor(P, N, c_loc(Q), c_loc(Cnt), c_loc(V)), &
               'uegccd_exchange_structure_factor')
    Write(82,*) "--------"
    Write(82,*) "Starting a new structure factor"
    Write(82,*) "--------"
    Do K = 1, N
      Write(82,*) Q(1,K), Q(2,K), Q(3,K), Cnt(K), V(K)
    End Do
    Flush(82)
  End Subroutine WriteExSF

End Module HEG
