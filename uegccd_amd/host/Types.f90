! Host-side data of a single-point calculation: the Input settings
! (UEGInfoType, source/Types.f90:29-82) and the basis (HEGDataType,
! source/Types.f90:7-27).  Only the fields the CCD path reads are kept.
Module Types
  Use Precision
  Implicit None

  Type UEGInfoType
    Integer :: NElectron = 0, NOcc = 0, NAO = 0, MaxKPoint = 0, CoreN = 0
    Integer :: NumRSPoints = 1, TACalcN = 0
    Integer :: IRange(15) = 0
    Real (Kind=pr) :: rSMin = 0, rSMax = 0, NTwist(3) = 0
    Real (Kind=pr) :: CoreFactor = 0, fcutMultiplier = 0, Nfermi = 0
    Real (Kind=pr) :: ExchangeFactor = 1, MadelungFactor = 1, Gap = 0, GapII = 0
    Real (Kind=pr) :: fcut = 0, corefcut = 1
    Logical :: SafeERI = .False., DoOnlyMP2Grid = .False., fcut2ecut = .False.
    Logical :: DoRing = .False., DoXRing = .False., DoLadder = .False., DoMosaic = .False.
    Logical :: DoCalcTACCD = .False., DodRPASOSEX = .False.
    Logical :: DoTruncCoulombHF = .False., DoTruncCoulombAll = .False., DoKS = .False.
    Logical :: DoSingleCalc = .False., DoSFCalcMP2 = .False., DoSFCalcCCD = .False.
    Logical :: DoSkipTA = .False.
  End Type UEGInfoType

  Type HEGDataType
    Integer :: Nel = 0
    Real (Kind=pr) :: L = 0, Omega = 0, Madelung = 0, EHF = 0, ECorr = 0, rS = 0
    Real (Kind=pr) :: ScreeningDistance = 0, DenomFactor = 2, TolMax = 1e-8_pr, FailRatio = 100
    Integer, Allocatable :: nVec(:,:)          ! (NAO,3)
    Real (Kind=pr), Allocatable :: Eigen(:)    ! (NAO)
  End Type HEGDataType

End Module Types
