"""Offline tile sweep for the tuned kernels (run on the GPU):

    python -m uegccd_amd.tune [--reps 40] [--out uegccd_amd/tuned_tiles.json]

For every representative system (N 14/38/54, cutoffs 5/9/16/25, the Gamma
point and a twist) it solves the system once, then times the contraction
kernel (ladder tiles, hole-hole units, ring items) with uegccd_kernel_time
for each ladder k-step group (1, 2, 4), in FP64 and in the FP32 phase of
mixed precision, and for each pair of k-steps per round of loads (ring
tiles 4/8, hole-hole tiles 4/8) in FP64.  Systems are grouped into classes by their largest pair-momentum
block (pairs <= 8/16/32/64) and its virtual list (<= 64/128/256/512/1024);
each class takes the choice with the lowest time summed over its systems,
each relative to its own default time.  The table is written as JSON; the
build compiles it into libuegccd (uegccd_amd/build.py: write_tuned_tiles).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(PKG))

from uegccd_amd import lib as L  # noqa: E402
from uegccd_amd.driver import CCDContext, init_heg, options  # noqa: E402
from uegccd_amd.inputfile import parse_input, standard_input  # noqa: E402

NOO_BOUNDS = (8, 16, 32, 64)
NVV_BOUNDS = (64, 128, 256, 512, 1024)
LADDER_GROUPS = (1, 2, 4)
CONTRACT_STEPS = ((4, 4), (4, 8), (8, 4), (8, 8))
DEFAULT = {"ladder_group": 2, "quad_steps": 8, "hh_steps": 8}
SYSTEMS = [(nel, cut, rs, tw) for nel, rs in ((14, 1.0), (38, 1.0), (54, 2.0)) for cut in (5, 9, 16, 25)
           for tw in ((0.0, 0.0, 0.0), (0.13, -0.27, 0.31))]


def bound(x, bounds):
    for b in bounds:
        if x <= b:
            return b
    return None


def kernel_us(lib, ctx, which, reps):
    us = ctypes.c_double()
    L.check(lib.uegccd_kernel_time(ctx.handle, which, reps, ctypes.byref(us)), "uegccd_kernel_time")
    return us.value


def set_tiles(lib, ctx, g, qs, hs):
    t = L.Tiles(g, qs, hs)
    L.check(lib.uegccd_set_tiles(ctx.handle, ctypes.byref(t)), "uegccd_set_tiles")


def sweep_system(lib, nel, cut, rs, tw, reps):
    info = parse_input(standard_input(nel, cut, rs, tw))
    heg = init_heg(info)
    out = {"nel": nel, "cutoff": cut, "rs": rs, "twist": list(tw), "nao": heg.NAO}
    for prec in ("fp64", "mixed"):
        ctx = CCDContext(heg, info, options(mode="diis", ndiis=8, precision=prec))
        _, _, tab, tba = ctx.mp2()
        ctx.solve(tab, tba)
        t, noo, nvv = L.Tiles(), ctypes.c_int32(), ctypes.c_int32()
        L.check(lib.uegccd_get_tiles(ctx.handle, ctypes.byref(t), ctypes.byref(noo), ctypes.byref(nvv)),
                "uegccd_get_tiles")
        out["noo_max"], out["nvv_max"] = noo.value, nvv.value
        lad = {}
        for g in LADDER_GROUPS:
            set_tiles(lib, ctx, g, 8, 8)
            lad[str(g)] = kernel_us(lib, ctx, 1, reps)
        out[f"ladder_us_{prec}"] = lad
        if prec == "fp64":
            con = {}
            for qs, hs in CONTRACT_STEPS:
                set_tiles(lib, ctx, 2, qs, hs)
                con[f"{qs},{hs}"] = kernel_us(lib, ctx, 1, reps)
            out["contract_us"] = con
        ctx.close()
    print(json.dumps(out), flush=True)
    return out


def choose(systems):
    classes = {}
    for s in systems:
        key = (bound(s["noo_max"], NOO_BOUNDS), bound(s["nvv_max"], NVV_BOUNDS))
        if None in key:
            continue
        classes.setdefault(key, []).append(s)
    table = []
    for (noo_b, nvv_b), members in sorted(classes.items()):
        con = min(CONTRACT_STEPS, key=lambda c: sum(m["contract_us"][f"{c[0]},{c[1]}"] / m["contract_us"]["8,8"]
                                                    for m in members))
        for prec in ("fp64", "mixed"):
            g = min(LADDER_GROUPS, key=lambda g: sum(m[f"ladder_us_{prec}"][str(g)] / m[f"ladder_us_{prec}"]["2"]
                                                     for m in members))
            table.append({"noo_max": noo_b, "nvv_max": nvv_b, "precision": prec, "ladder_group": g,
                          "quad_steps": con[0], "hh_steps": con[1],
                          "systems": [f"N={m['nel']} cutoff {m['cutoff']} twist {m['twist']}" for m in members]})
    return table


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=40)
    ap.add_argument("--out", default=os.path.join(PKG, "tuned_tiles.json"))
    args = ap.parse_args(argv)
    import torch  # before the library takes the device
    device = torch.cuda.get_device_name(0)
    lib = L.load()
    systems = [sweep_system(lib, *s, args.reps) for s in SYSTEMS]
    doc = {"command": "python -m uegccd_amd.tune", "device": device,
           "reps": args.reps, "default": DEFAULT, "classes": choose(systems), "measurements": systems}
    with open(args.out, "w") as f:
        json.dump(doc, f, indent=1)
    print(f"wrote {args.out}: {len(doc['classes'])} classes")


if __name__ == "__main__":
    main()
