"""Extract the golden vectors the reference's stored outputs hold for the CCD
path into tests/golden/reference_fixtures.json.

Run in a container where the reference tree is mounted (default
/root/reference); the JSON it writes is what the tests read, so nothing reads
the reference at test time.  Only numbers are taken (inputs and printed
results), never source text.

Sources (paths relative to the reference root):
  source/Test/Output      N=24 r_s=5 single point (E(SCF), eigenvalues, E(2),
                          the 70-iteration CCD trajectory, Ec, MaxRes), the 10
                          twist-averaging MP2 blocks and the final TA CCD
  source/Test/fort.60     basis n-vectors / eigenvalues at r_s=1 and 5, Madelung
  source/Test/fort.59     the 10 twists, averaged eigenvalues, special twist
  benchmarks/ZRun-dRPA/Output, fort.60      N=14 cutoff 5 (57 orbitals): MP2 and dRPA/SOSEX
  benchmarks/ZRun-v6-N14/fort.59, Output    N=14 cutoff 2, 100 twists
  benchmarks/ZRun-dRPA/old-dRPA/*-data      N=14 r_s=1 dRPA/SOSEX basis scan, cutoffs
                          5..665 (cutoff, Ec(dRPA), Ec(SOSEX), a run time, 2M)
"""
from __future__ import annotations

import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
FLOAT = r"[-+]?\d*\.\d+(?:[EeDd][-+]?\d+)?"


def read(root, rel):
    with open(os.path.join(root, rel)) as f:
        return f.read().splitlines()


def f(x):
    return float(x.replace("D", "E").replace("d", "e"))


def output_blocks(lines):
    """Split an Output file into MP2 and CCD blocks (in file order)."""
    blocks = []
    cur = None
    for ln in lines:
        if "E(SCF) is" in ln and "+" not in ln:
            cur = {"kind": "mp2", "ehf": f(re.findall(FLOAT, ln)[0]), "eig": []}
            blocks.append(cur)
        elif cur is not None and cur["kind"] == "mp2" and re.match(r"^\s+\*\s+" + FLOAT + r"\s+\*\s*$", ln):
            cur["eig"].append(f(re.findall(FLOAT, ln)[0]))
        elif "E(SCF) + E(2) is" in ln:
            cur["ehf_plus_e2"] = f(re.findall(FLOAT, ln)[0])
        elif "E(2) is" in ln:
            cur["e2"] = f(re.findall(FLOAT, ln)[0])
        elif "CCD summary follows" in ln:
            cur = {"kind": "ccd", "trace": []}
            blocks.append(cur)
        elif cur is not None and cur["kind"] == "ccd":
            m = re.match(r"^\s+\*\s+(" + FLOAT + r")\s+(\d+)\s+(" + FLOAT + r")\s+\*\s*$", ln)
            if m:
                cur["trace"].append([f(m.group(1)), int(m.group(2)), f(m.group(3))])
            elif "converged in" in ln:
                cur["niter"] = int(re.findall(r"in\s+(\d+)", ln)[0])
            elif "Ec(CCD) is" in ln or "Ec(dRPA) is" in ln:
                cur["ecorr"] = f(re.findall(FLOAT, ln)[0])
            elif "Ec(SOSEX) is" in ln:
                cur["esosex"] = f(re.findall(FLOAT, ln)[0])
            elif "Max CCD Residual" in ln:
                cur["maxres"] = f(re.findall(FLOAT, ln)[0])
    return blocks


def fort60_tables(lines):
    """Basis listings 'p n1 n2 n3 eigen' and Madelung values, in order."""
    tables, cur, mads, Ls = [], None, [], []
    for ln in lines:
        t = ln.split()
        if len(t) == 5 and all(re.match(r"^-?\d+$", x) for x in t[:4]):
            if cur is None or int(t[0]) == 1:
                cur = []
                tables.append(cur)
            cur.append([int(t[1]), int(t[2]), int(t[3]), f(t[4])])
        else:
            cur = None if len(t) != 5 else cur
        if "Madelung:" in ln or "Madelung constant calculated as" in ln:
            mads.append(f(t[-1]))
        if ln.strip().startswith("L:"):
            Ls.append(f(t[-1]))
    return tables, mads, Ls


def main(root="/root/reference"):
    out = {}
    # --- N=24, r_s=5, twist (0.1,0.2,0.45), cutoff 40, fcut 2 (source/Test) ---
    blk = output_blocks(read(root, "source/Test/Output"))
    mp2s = [b for b in blk if b["kind"] == "mp2"]
    ccds = [b for b in blk if b["kind"] == "ccd"]
    tables, mads, Ls = fort60_tables(read(root, "source/Test/fort.60"))
    f59 = [ln.split() for ln in read(root, "source/Test/fort.59")]
    twists = [[f(x) for x in row[:3]] for row in f59[:10]]
    avg_eig = [f(x) for x in f59[13] if re.match("^" + FLOAT + "$", x)]
    special = [f(x) for x in f59[-2][-3:]]
    out["n24"] = {
        "params": {"nel": 24, "maxkpoint": 40, "rs": 5.0, "twist": [0.1, 0.2, 0.45],
                   "fcut": 2.0, "corefcut": 1.0},
        "ehf": mp2s[0]["ehf"], "eigen": mp2s[0]["eig"], "e2": mp2s[0]["e2"],
        "ccd_trace": ccds[0]["trace"], "ccd_niter": ccds[0]["niter"],
        "ccd_ecorr": ccds[0]["ecorr"], "ccd_maxres": ccds[0]["maxres"],
        "basis_rs1": tables[0], "basis_rs5": tables[1],
        "madelung_rs1": mads[1], "madelung_rs5": mads[3], "L_rs1": Ls[0], "L_rs5": Ls[1],
        "ta_twists": twists,
        "ta_mp2": [{"ehf": b["ehf"], "e2": b["e2"], "eig": b["eig"]} for b in mp2s[1:11]],
        "ta_fort59_hf_mp2": [[f(row[3]), f(row[4])] for row in f59[:10]],
        "ta_special_twist": special,
        "ta_avg_eigen": avg_eig,
        "ta_final_e2": mp2s[11]["e2"],
        "ta_final_ehf": mp2s[11]["ehf"],
        "ta_final_trace": ccds[1]["trace"], "ta_final_niter": ccds[1]["niter"],
        "ta_final_ecorr": ccds[1]["ecorr"], "ta_final_maxres": ccds[1]["maxres"],
        # the whole fort.59 of the twist-averaging run, as printed (per-twist
        # E_HF/E(2)/E_c, averages, squared averages, errors, eigenvalue and
        # connectivity averages, connectivity distances, the special twist)
        "ta_fort59_lines": read(root, "source/Test/fort.59"),
    }
    # --- N=14, r_s=1, Gamma, cutoff 5 (benchmarks/ZRun-dRPA) ---
    blk = output_blocks(read(root, "benchmarks/ZRun-dRPA/Output"))
    tables, mads, Ls = fort60_tables(read(root, "benchmarks/ZRun-dRPA/fort.60"))
    out["n14_cut5"] = {
        "params": {"nel": 14, "maxkpoint": 5, "rs": 1.0, "twist": [0.0, 0.0, 0.0]},
        "ehf": blk[0]["ehf"], "eigen": blk[0]["eig"], "e2": blk[0]["e2"],
        "madelung": mads[0] if mads else None, "basis": tables[0] if tables else None,
        # the dRPA/SOSEX run of that Input (DodRPASOSEX T, mosaics off): the
        # printed trajectory (E, NIter, dT), Ec(dRPA), Ec(SOSEX), max residual
        "drpa_trace": blk[1]["trace"], "drpa_niter": blk[1]["niter"],
        "drpa_ecorr": blk[1]["ecorr"], "drpa_esosex": blk[1]["esosex"],
        "drpa_maxres": blk[1]["maxres"],
    }
    # --- N=14, r_s=1, cutoff 2, first twist of the v6 run (benchmarks/ZRun-v6-N14) ---
    rows = [ln.split() for ln in read(root, "benchmarks/ZRun-v6-N14/fort.59")]
    blk = output_blocks(read(root, "benchmarks/ZRun-v6-N14/Output"))
    out["v6_twist1"] = {
        "params": {"nel": 14, "maxkpoint": 2, "rs": 1.0,
                   "twist": [f(x) for x in rows[0][:3]], "safe_eri": False},
        "ehf": f(rows[0][3]), "e2": f(rows[0][4]), "ccd_ecorr": f(rows[0][5]),
        "eigen": blk[0]["eig"], "ccd_trace": blk[1]["trace"], "ccd_niter": blk[1]["niter"],
        "ccd_maxres": blk[1]["maxres"],
    }
    # --- N=14, r_s=1, Gamma dRPA/SOSEX scan over the cutoff (benchmarks/ZRun-dRPA/old-dRPA):
    # three runs (high/med/low) whose energies agree row for row (the 2M
    # column of the low run reads 0 in its cutoff-610 row); the fourth column
    # (a run time) is left out ---
    scans = [[ln.split() for ln in read(root, f"benchmarks/ZRun-dRPA/old-dRPA/{k}-1.0-data") if ln.strip()]
             for k in ("high", "med", "low")]
    for other in scans[1:]:
        assert [r[:3] for r in other] == [r[:3] for r in scans[0]][:len(other)]
    # (the Input has no fcut key, and 2M is the whole cutoff basis on every
    # row: fcut 1000, 221,359 orbitals for N=14, truncates none of them)
    out["drpa_scan"] = {
        "params": {"nel": 14, "rs": 1.0, "twist": [0.0, 0.0, 0.0], "do_mosaic": False, "fcut": 1000.0},
        "rows": [[int(r[0]), f(r[1]), f(r[2]), int(r[4])] for r in scans[0]],
        "columns": ["cutoff", "ecorr_drpa", "ecorr_sosex", "two_m"],
    }
    path = os.path.join(HERE, "reference_fixtures.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main(*sys.argv[1:])
