"""The reference's ``Input`` file format (source/IO.f90:14-391), unchanged.

``Keyword: Value`` lines, ``!`` starts a comment, each of the 44 keywords must
appear exactly once (``# Electrons`` and ``Momentum Cutoff`` may be omitted),
``NTwist`` is optional, an unknown keyword stops parsing with an error.  The
error texts are the reference's ``Stop`` / format-1000..1020 messages.
"""
from __future__ import annotations

from dataclasses import dataclass, field

LLINE = 79  # characters read per line (source/IO.f90:20,70)

PARAM_NAMES = (
    "# Electrons", "Momentum Cutoff", "rSMin", "rSMax", "# rS Points", "Safe ERI",
    "Do Rings", "Do XRings", "Do Ladders", "Do Mosaics",
    "Rings Range", "XRings Range", "Ladders Range", "Mosaics Range", "DriverDir Range",
    "DriverEx Range", "Energy Range", "LinRings Range", "QuadRings Range", "DRings Range",
    "ExRings Range", "LinLadd Range", "QuadLadd Range", "DLadders Range", "ExLadders Range",
    "CoreFactor", "TACalcN", "DoCalcTACCD", "DodRPASOSEX", "DoTruncCoulombHF",
    "DoTruncCoulombAll", "DoKS", "MadelungFactor", "DoSingleCalc", "Gap", "fcut", "corefcut",
    "fcut2ecut", "GapII", "DoSFCalcMP2", "DoSFCalcCCD", "DoSkipTA", "DoOnlyMP2Grid",
    "EndInput",
)
RANGE_NAMES = PARAM_NAMES[10:25]


class InputError(ValueError):
    """A condition on which the reference's ReadInput stops."""


@dataclass
class UEGInfo:
    """The fields of UEGInfoType (source/Types.f90:29-82) that ReadInput sets."""
    NElectron: int = 0
    NOcc: int = 0
    MaxKPoint: int = 0
    CoreN: int = 0
    NumRSPoints: int = 1
    irange: list = field(default_factory=lambda: [0] * 15)
    TACalcN: int = 0
    rSMin: float = 0.0
    rSMax: float = 0.0
    NTwist: tuple = (0.0, 0.0, 0.0)
    CoreFactor: float = 0.0
    fcutMultiplier: float = 0.0
    Nfermi: float = 0.0
    exchangefactor: float = 1.0
    madelungfactor: float = 1.0
    Gap: float = 0.0
    GapII: float = 0.0
    fcut: float = 0.0
    corefcut: float = 1.0
    SafeERI: bool = False
    DoOnlyMP2Grid: bool = False
    fcut2ecut: bool = False
    DoRing: bool = False
    DoXRing: bool = False
    DoLadder: bool = False
    DoMosaic: bool = False
    DoCalcTACCD: bool = False
    DodRPASOSEX: bool = False
    DoTruncCoulombHF: bool = False
    DoTruncCoulombAll: bool = False
    DoKS: bool = False
    DoSingleCalc: bool = False
    DoSFCalcMP2: bool = False
    DoSFCalcCCD: bool = False
    DoSkipTA: bool = False
    NAO: int = 0  # set by the basis construction
    # Not an Input keyword (the reference cuts on the untwisted |n|^2 <= MaxKPoint,
    # source/HEG.f90:190-194): callers opt in to the twisted sphere
    # |n + twist|^2 <= MaxKPoint by setting it.
    CutoffTwisted: bool = False

    @property
    def rs(self) -> float:
        return self.rSMin


def _first(value: str) -> str:
    tok = value.replace(",", " ").split()
    if not tok:
        raise InputError("Error in Input")
    return tok[0]


def _int(value: str) -> int:
    try:
        return int(_first(value))
    except ValueError:
        raise InputError("Error in Input") from None


def _real(value: str) -> float:
    try:
        return float(_first(value).replace("d", "e").replace("D", "e"))
    except ValueError:
        raise InputError("Error in Input") from None


def _logical(value: str) -> bool:
    t = _first(value).lstrip(".")
    if t[:1] in ("T", "t"):
        return True
    if t[:1] in ("F", "f"):
        return False
    raise InputError("Error in Input")


def _parse_line(line: str):
    """ParseLine (source/IO.f90:355-391)."""
    work = line
    cpos = work.find("!")
    if cpos >= 0:
        work = work[:cpos]
    work = "".join(" " if ord(ch) <= 31 else ch for ch in work)
    if not work.strip():
        return "SKIP", ""
    colon = work.find(":")
    if colon < 0:
        return work, ""
    return work[:colon], work[colon + 1:]


def parse_input(text: str) -> UEGInfo:
    """ReadInput (source/IO.f90:14-348) on the text of an Input file."""
    info = UEGInfo()
    set_once = {k: False for k in PARAM_NAMES}
    set_twice = {k: False for k in PARAM_NAMES}
    messages = []
    unknown_line = 0
    for lineno, raw in enumerate(text.splitlines(), start=1):
        key, value = _parse_line(raw[:LLINE])
        key = key.strip()
        if key in set_once:
            if set_once[key]:
                set_twice[key] = True
                messages.append(f"Error: Parameter {key} multiply set.")
            set_once[key] = True
        if key == "SKIP":
            continue
        if key == "# Electrons":
            info.NElectron = _int(value)
            info.NOcc = info.NElectron // 2
            if info.NElectron < 0:
                raise InputError("Must have positive number of electrons")
            if info.NElectron % 2 == 1:
                raise InputError("Must have closed shell")
        elif key == "Momentum Cutoff":
            info.MaxKPoint = _int(value)
        elif key == "rSMin":
            info.rSMin = _real(value)
        elif key == "rSMax":
            info.rSMax = _real(value)
        elif key == "# rS Points":
            info.NumRSPoints = _int(value)
            if info.NumRSPoints > 1:
                raise InputError("Cannot have more than one rs point")
        elif key == "Safe ERI":
            info.SafeERI = _logical(value)
        elif key == "Do Rings":
            info.DoRing = _logical(value)
        elif key == "Do XRings":
            info.DoXRing = _logical(value)
        elif key == "Do Ladders":
            info.DoLadder = _logical(value)
        elif key == "Do Mosaics":
            info.DoMosaic = _logical(value)
        elif key in RANGE_NAMES:
            info.irange[RANGE_NAMES.index(key)] = _int(value)
        elif key == "NTwist":
            tok = value.replace(",", " ").split()
            if len(tok) < 3:
                raise InputError("Error in Input")
            info.NTwist = tuple(float(t.replace("d", "e").replace("D", "e")) for t in tok[:3])
        elif key == "CoreFactor":
            info.CoreFactor = _real(value)
            info.CoreN = int((info.NElectron // 2) * info.CoreFactor)
        elif key == "DoCalcTACCD":
            info.DoCalcTACCD = _logical(value)
        elif key == "TACalcN":
            info.TACalcN = _int(value)
        elif key == "DodRPASOSEX":
            info.DodRPASOSEX = _logical(value)
        elif key == "DoTruncCoulombHF":
            info.DoTruncCoulombHF = _logical(value)
            if info.DoTruncCoulombHF:
                raise InputError("Not Implemented")
        elif key == "DoTruncCoulombAll":
            info.DoTruncCoulombAll = _logical(value)
            if info.DoTruncCoulombAll:
                raise InputError("Not Implemented")
        elif key == "DoKS":
            info.DoKS = _logical(value)
            info.exchangefactor = 0.0 if info.DoKS else 1.0
        elif key == "MadelungFactor":
            info.madelungfactor = _real(value)
        elif key == "DoSingleCalc":
            info.DoSingleCalc = _logical(value)
        elif key == "DoSFCalcMP2":
            info.DoSFCalcMP2 = _logical(value)
        elif key == "DoSFCalcCCD":
            info.DoSFCalcCCD = _logical(value)
        elif key == "DoSkipTA":
            info.DoSkipTA = _logical(value)
        elif key == "DoOnlyMP2Grid":
            info.DoOnlyMP2Grid = _logical(value)
        elif key == "Gap":
            info.Gap = _real(value)
        elif key == "fcut":
            info.fcut = _real(value)
        elif key == "corefcut":
            info.corefcut = _real(value)
        elif key == "fcut2ecut":
            info.fcut2ecut = _logical(value)
            if info.fcut2ecut:
                # source/IO.f90:291-301
                info.Nfermi = info.NElectron ** (2.0 / 3.0) / 4.0
                info.fcutMultiplier = info.fcut
                if info.fcut < 2.0:
                    info.fcutMultiplier = 2.0
                info.MaxKPoint = int(info.Nfermi * info.fcutMultiplier * 2.0)
                if info.Nfermi < 10:
                    info.MaxKPoint = int(10 * info.fcutMultiplier * 2.0)
        elif key == "GapII":
            info.GapII = _real(value)
        elif key == "EndInput":
            if not _logical(value):
                raise InputError("EndInput must be T")
        else:
            unknown_line = lineno
            break
    set_once["# Electrons"] = True
    set_once["Momentum Cutoff"] = True
    if unknown_line:
        messages.append(f"Error: Line number {unknown_line:4d} of Input not recognized; "
                        "subsequent lines not read.")
    for k in PARAM_NAMES:
        if not set_once[k]:
            messages.append(f"Error: Parameter {k} not set.")
    if unknown_line or any(set_twice.values()) or not all(set_once.values()):
        raise InputError("Error in Input\n" + "\n".join(messages))
    return info


def read_input(path: str) -> UEGInfo:
    if not __import__("os").path.exists(path):
        raise InputError("No Input file")
    with open(path) as f:
        return parse_input(f.read())


def write_input(info_values: dict) -> str:
    """An Input text with every keyword, from a {keyword: value} mapping."""
    lines = []
    for k in PARAM_NAMES:
        v = info_values[k]
        if isinstance(v, bool):
            v = "T" if v else "F"
        lines.append(f"{k + ':':<19}{v}")
    if "NTwist" in info_values:
        t = info_values["NTwist"]
        lines.insert(25, f"{'NTwist:':<19}{t[0]!r} {t[1]!r} {t[2]!r}")
    return "\n".join(lines) + "\n"


def standard_input(nel: int, cutoff: int, rs: float, twist=(0.0, 0.0, 0.0), **over) -> str:
    """A single-point CCD Input with every term on and every range 0."""
    vals = {k: 0 for k in RANGE_NAMES}
    vals.update({
        "# Electrons": nel, "Momentum Cutoff": cutoff, "rSMin": rs, "rSMax": rs,
        "# rS Points": 1, "Safe ERI": True, "Do Rings": True, "Do XRings": True,
        "Do Ladders": True, "Do Mosaics": True, "CoreFactor": 0.0, "TACalcN": 0,
        "DoCalcTACCD": False, "DodRPASOSEX": False, "DoTruncCoulombHF": False,
        "DoTruncCoulombAll": False, "DoKS": False, "MadelungFactor": 1.0,
        "DoSingleCalc": True, "Gap": 0.0, "fcut": 100.0, "corefcut": 1,
        "fcut2ecut": False, "GapII": -0.01, "DoSFCalcMP2": False, "DoSFCalcCCD": False,
        "DoSkipTA": True, "DoOnlyMP2Grid": False, "EndInput": True, "NTwist": tuple(twist),
    })
    vals.update(over)
    return write_input(vals)
