! The Input file reader: the reference's format and checks (source/IO.f90:14-391).
! 44 keywords, each exactly once ('# Electrons' and 'Momentum Cutoff' may be
! omitted), 'NTwist' optional, '!' comments, unknown keyword = error.
Module IO
  Use Precision
  Use Types, Only: UEGInfoType
  Implicit None
  Private
  Public :: ReadInput

  Integer, Parameter :: NParams = 44, LLine = 79

Contains

  Subroutine ReadInput(UEGInfo)
    Type (UEGInfoType), Intent(Out) :: UEGInfo
    Character (Len=19), Parameter :: Names(NParams) = [Character(Len=19) :: &
      '# Electrons', 'Momentum Cutoff', 'rSMin', 'rSMax', '# rS Points', 'Safe ERI', &
      'Do Rings', 'Do XRings', 'Do Ladders', 'Do Mosaics', 'Rings Range', 'XRings Range', &
      'Ladders Range', 'Mosaics Range', 'DriverDir Range', 'DriverEx Range', 'Energy Range', &
      'LinRings Range', 'QuadRings Range', 'DRings Range', 'ExRings Range', 'LinLadd Range', &
      'QuadLadd Range', 'DLadders Range', 'ExLadders Range', 'CoreFactor', 'TACalcN', &
      'DoCalcTACCD', 'DodRPASOSEX', 'DoTruncCoulombHF', 'DoTruncCoulombAll', 'DoKS', &
      'MadelungFactor', 'DoSingleCalc', 'Gap', 'fcut', 'corefcut', 'fcut2ecut', 'GapII', &
      'DoSFCalcMP2', 'DoSFCalcCCD', 'DoSkipTA', 'DoOnlyMP2Grid', 'EndInput']
    Character (Len=LLine) :: Line, Key, Value
    Logical :: SetOnce(NParams), SetTwice(NParams), Exists, EndFlag, Unknown
    Integer :: I, IOS, LineNumber, BadLine

    Write(6,*) 'Reading Input...'
    Inquire(File='Input', Exist=Exists)
    If (.not. Exists) Stop 'No Input file'
    SetOnce = .False.
    SetTwice = .False.
    Unknown = .False.
    BadLine = 0
    LineNumber = 0
    Open(4, File='Input', Status='Old')
    Do
      LineNumber = LineNumber + 1
      Read(4, '(A)', IOStat=IOS) Line
      If (IOS /= 0) Exit
      Call ParseLine(Line, Key, Value)
      Do I = 1, NParams
        If (Trim(Names(I)) == Trim(AdjustL(Key))) Then
          If (SetOnce(I)) SetTwice(I) = .True.
          If (SetTwice(I)) Write(6,'("Error: Parameter ",A," multiply set.")') Trim(Names(I))
          SetOnce(I) = .True.
        End If
      End Do
      Select Case (Trim(AdjustL(Key)))
        Case ('SKIP')
        Case ('# Electrons')
          Read(Value,*) UEGInfo%NElectron
          UEGInfo%NOcc = UEGInfo%NElectron/2
          If (UEGInfo%NElectron < 0) Stop 'Must have positive number of electrons'
          If (Mod(UEGInfo%NElectron,2) == 1) Stop 'Must have closed shell'
        Case ('Momentum Cutoff');  Read(Value,*) UEGInfo%MaxKPoint
        Case ('rSMin');            Read(Value,*) UEGInfo%rSMin
        Case ('rSMax');            Read(Value,*) UEGInfo%rSMax
        Case ('# rS Points')
          Read(Value,*) UEGInfo%NumRSPoints
          If (UEGInfo%NumRSPoints > 1) Stop 'Cannot have more than one rs point'
        Case ('Safe ERI');         Read(Value,*) UEGInfo%SafeERI
        Case ('Do Rings');         Read(Value,*) UEGInfo%DoRing
        Case ('Do XRings');        Read(Value,*) UEGInfo%DoXRing
        Case ('Do Ladders');       Read(Value,*) UEGInfo%DoLadder
        Case ('Do Mosaics');       Read(Value,*) UEGInfo%DoMosaic
        Case ('Rings Range', 'XRings Range', 'Ladders Range', 'Mosaics Range', &
              'DriverDir Range', 'DriverEx Range', 'Energy Range', 'LinRings Range', &
              'QuadRings Range', 'DRings Range', 'ExRings Range', 'LinLadd Range', &
              'QuadLadd Range', 'DLadders Range', 'ExLadders Range')
          Do I = 11, 25
            If (Trim(Names(I)) == Trim(AdjustL(Key))) Read(Value,*) UEGInfo%IRange(I-10)
          End Do
        Case ('NTwist');           Read(Value,*) UEGInfo%NTwist
        Case ('CoreFactor')
          Read(Value,*) UEGInfo%CoreFactor
          UEGInfo%CoreN = Int(UEGInfo%NElectron/2*UEGInfo%CoreFactor)
        Case ('DoCalcTACCD');      Read(Value,*) UEGInfo%DoCalcTACCD
        Case ('TACalcN');          Read(Value,*) UEGInfo%TACalcN
        Case ('DodRPASOSEX');      Read(Value,*) UEGInfo%DodRPASOSEX
        Case ('DoTruncCoulombHF')
          Read(Value,*) UEGInfo%DoTruncCoulombHF
          If (UEGInfo%DoTruncCoulombHF) Stop 'Not Implemented'
        Case ('DoTruncCoulombAll')
          Read(Value,*) UEGInfo%DoTruncCoulombAll
          If (UEGInfo%DoTruncCoulombAll) Stop 'Not Implemented'
        Case ('DoKS')
          Read(Value,*) UEGInfo%DoKS
          UEGInfo%ExchangeFactor = Merge(0.0_pr, 1.0_pr, UEGInfo%DoKS)
        Case ('MadelungFactor');   Read(Value,*) UEGInfo%MadelungFactor
        Case ('DoSingleCalc');     Read(Value,*) UEGInfo%DoSingleCalc
        Case ('DoSFCalcMP2');      Read(Value,*) UEGInfo%DoSFCalcMP2
        Case ('DoSFCalcCCD');      Read(Value,*) UEGInfo%DoSFCalcCCD
        Case ('DoSkipTA');         Read(Value,*) UEGInfo%DoSkipTA
        Case ('DoOnlyMP2Grid');    Read(Value,*) UEGInfo%DoOnlyMP2Grid
        Case ('Gap');              Read(Value,*) UEGInfo%Gap
        Case ('fcut');             Read(Value,*) UEGInfo%fcut
        Case ('corefcut');         Read(Value,*) UEGInfo%corefcut
        Case ('fcut2ecut')
          Read(Value,*) UEGInfo%fcut2ecut
          If (UEGInfo%fcut2ecut) Then
            UEGInfo%Nfermi = UEGInfo%NElectron**(2.0_pr/3.0_pr)/4.0_pr
            UEGInfo%fcutMultiplier = UEGInfo%fcut
            If (UEGInfo%fcut < 2.0_pr) UEGInfo%fcutMultiplier = 2.0_pr
            UEGInfo%MaxKPoint = Int(UEGInfo%Nfermi*UEGInfo%fcutMultiplier*2.0_pr)
            If (UEGInfo%Nfermi < 10) UEGInfo%MaxKPoint = Int(10*UEGInfo%fcutMultiplier*2.0_pr)
          End If
        Case ('GapII');            Read(Value,*) UEGInfo%GapII
        Case ('EndInput')
          Read(Value,*) EndFlag
          If (.not. EndFlag) Stop 'EndInput must be T'
        Case Default
          Unknown = .True.
          BadLine = LineNumber
          Write(6,*) Trim(Value)
          Exit
      End Select
    End Do
    Close(4)
    SetOnce(1:2) = .True.
    If (Unknown) Write(6,'("Error: Line number ",I4," of Input not recognized; ", &
                         & "subsequent lines not read.")') BadLine
    Do I = 1, NParams
      If (.not. SetOnce(I)) Write(6,'("Error: Parameter ",A," not set.")') Trim(Names(I))
    End Do
    If (Unknown .or. Any(SetTwice) .or. .not. All(SetOnce)) Stop 'Error in Input'
  End Subroutine ReadInput

  ! Comment stripping and 'Keyword: Value' split (source/IO.f90:355-391).
  Subroutine ParseLine(Line, Key, Value)
    Character (Len=*), Intent(In) :: Line
    Character (Len=*), Intent(Out) :: Key, Value
    Character (Len=Len(Line)) :: W
    Integer :: I, P
    W = Line
    P = Index(W, '!')
    If (P /= 0) W = W(1:P-1)
    Do I = 1, Len(W)
      If (IAChar(W(I:I)) <= 31) W(I:I) = ' '
    End Do
    If (Len_Trim(W) == 0) Then
      Key = 'SKIP'
      Value = ' '
      Return
    End If
    P = Index(W, ':')
    Key = W(1:Max(P-1,0))
    Value = W(P+1:)
  End Subroutine ParseLine

End Module IO
