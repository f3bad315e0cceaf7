"""Build recipe for libuegccd.so (HIP, gfx950), the ZCode Fortran host and the
CPU oracle used by the tests.

    python -m uegccd_amd.build            # everything
    python -m uegccd_amd.build lib        # only libuegccd.so

Objects go to uegccd_amd/_build/ (git-ignored); the shared library and the
ZCode executable are written in-tree so they travel with the repository.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libuegccd.so")
ZCODE = os.path.join(HOST, "ZCode")

HIP_SOURCES = ["kernels.hip", "api.hip"]
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # no packed-FP32 (v_pk_*_f32) instructions: some of their forms have given
    # wrong results on these MI355X boxes
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
# Fortran host: the reference's file set (Main, IO, HEG-facing types, CCD,
# MP2, DIISCC, Wrappers) compiled in dependency order.
FORTRAN_SOURCES = ["Precision.f90", "Types.f90", "Wrappers.f90", "IO.f90", "HEG.f90",
                   "DIISCC.f90", "CCD.f90", "dRPA.f90", "MP2.f90", "TwistAverage.f90", "Main.f90"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        out = (r.stdout + r.stderr).strip()
        raise RuntimeError(f"build step failed ({' '.join(cmd[:3])} ...):\n{out}")
    return r


def hipcc():
    for cand in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


TUNED_JSON = os.path.join(PKG, "tuned_tiles.json")
TUNED_INC = os.path.join(CSRC, "tuned_tiles.inc")


def write_tuned_tiles():
    """The tile table of the offline sweep (tuned_tiles.json, written by
    `python -m uegccd_amd.tune`) as the C initialisers api.hip includes; an
    empty table (every class takes the defaults) before the first sweep."""
    import json
    rows = []
    if os.path.exists(TUNED_JSON):
        for e in json.load(open(TUNED_JSON))["classes"]:
            rows.append("    {%d, %d, %d, %d, %d, %d}," % (
                e["noo_max"], e["nvv_max"], 1 if e["precision"] == "mixed" else 0,
                e["ladder_group"], e["quad_steps"], e["hh_steps"]))
    text = "// generated by uegccd_amd/build.py from tuned_tiles.json\n" + "\n".join(rows) + "\n"
    if not os.path.exists(TUNED_INC) or open(TUNED_INC).read() != text:
        with open(TUNED_INC, "w") as f:
            f.write(text)


def build_lib():
    os.makedirs(BUILD, exist_ok=True)
    write_tuned_tiles()
    objs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([hipcc(), *HIPCC_FLAGS, "-c", path, "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_zcode():
    flang = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(flang):
        raise RuntimeError("amdflang not found")
    if not os.path.isdir(HOST):
        return None
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in FORTRAN_SOURCES:
        path = os.path.join(HOST, src)
        if not os.path.exists(path):
            return None
        obj = os.path.join(BUILD, src.replace(".f90", ".o"))
        _run([flang, "-O2", "-c", path, "-o", obj, "-module-dir", BUILD, "-J", BUILD], cwd=BUILD)
        objs.append(obj)
    tmp = ZCODE + ".tmp"
    _run([flang, *objs, "-L" + PKG, "-luegccd", "-Wl,-rpath,$ORIGIN/..", "-o", tmp])
    os.replace(tmp, ZCODE)
    return ZCODE


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def main(argv):
    what = argv[1] if len(argv) > 1 else "all"
    if what in ("all", "lib"):
        print("built", build_lib())
    if what in ("all", "oracle"):
        build_oracle()
        print("built oracle")
    if what in ("all", "zcode"):
        z = build_zcode()
        print("built", z)


if __name__ == "__main__":
    main(sys.argv)
