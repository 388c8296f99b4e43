#!/usr/bin/env python3
"""Accuracy of the population storage forms against an FP64 run of the same lattice.

Runs the uniform periodic TGV (default 128^3, Re 1600, u0 0.05, to t* = 12) in FP64, FP32, FP16 and
BF16 storage on one GPU through the C ABI, samples energy, enstrophy and mass every --every steps, and
reports per form: the enstrophy-peak time and its offset from FP64, the maximum relative energy error,
the maximum relative enstrophy error, and the maximum mass drift difference. Optionally (--amr N0) the
same for the 3-level AMR with sensor refinement. Writes JSON (curves + summary) and prints a table.
"""
import argparse
import json
import math
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import numpy as np  # noqa: E402

import jpabi  # noqa: E402

FORMS = ["fp64", "fp32", "fp16", "bf16"]


def run_uniform(n, re, u0, tstar, every, prec):
    u = jpabi.Uniform(n, re=re, u0=u0, prec=jpabi.PREC_NAMES[prec])
    u.init()
    d = u.derived()
    steps = int(tstar / (u0 * d["kf"]) + 0.5)
    curve = []
    s = 0
    t0 = time.time()
    while True:
        E, ens, m = u.diag()
        curve.append((s * u0 * d["kf"], E, ens, m))
        if s >= steps:
            break
        k = min(every, steps - s)
        u.step(k)
        s += k
    wall = time.time() - t0
    u.close()
    return {"steps": steps, "curve": curve, "wall_s": wall}


def sensors_by_pos(a, n0):
    """The sensors the last init/regrid used: level 0 per coarse tile, level 1 scattered by position."""
    ng1 = (n0 // 4) ** 3
    s1 = np.zeros(ng1, np.float32)
    pos = a.table("f1pos")
    if pos.size:
        s1[pos] = a.table("sen1")
    return a.table("sen0"), s1


def run_amr(n0, re, u0, tstar, every, prec, adapt, script=None):
    """Dynamic sensor AMR. The FP64 run records the sensor fields of its init and of every regrid
    (script=None); the other forms replay them, so all forms advance on the same sequence of grids
    and the comparison measures storage error, not a different refinement history (threshold ties
    of the symmetric TGV are otherwise broken differently by each form's rounding)."""
    a = jpabi.Amr(n0, re=re, u0=u0, l1=jpabi.MODE_SENSOR, l2=jpabi.MODE_SENSOR, frac1=0.3, frac2=0.1, adapt=adapt,
                  prec=jpabi.PREC_NAMES[prec])
    rec = []
    if script is None:
        a.init()
        rec.append(sensors_by_pos(a, n0))
    else:
        a.init(*script[0])
    d = a.derived()
    steps = int(tstar / (u0 * d["kf"]) + 0.5)
    curve = []
    t0 = time.time()
    for s in range(steps + 1):
        if s > 0 and s % adapt == 0:
            if script is None:
                a.regrid()
                rec.append(sensors_by_pos(a, n0))
            else:
                a.regrid(*script[s // adapt])  # entry 0 is the init's
        if s % every == 0:
            E, ens, m = a.diag()
            curve.append((s * u0 * d["kf"], E, ens, m))
        if s < steps:
            a.step(1)
    wall = time.time() - t0
    c = a.counts()
    a.close()
    return {"steps": steps, "curve": curve, "wall_s": wall, "counts": c, "script": rec}


def summarise(runs):
    ref = runs["fp64"]["curve"]
    tref = [c[0] for c in ref]
    ipk = max(range(len(ref)), key=lambda i: ref[i][2])
    out = {}
    for f, r in runs.items():
        cur = r["curve"]
        assert [c[0] for c in cur] == tref
        ip = max(range(len(cur)), key=lambda i: cur[i][2])
        E0 = cur[0][1]
        eE = max(abs(c[1] - q[1]) / q[1] for c, q in zip(cur, ref))
        eZ = max(abs(c[2] - q[2]) / q[2] for c, q in zip(cur, ref))
        m0, q0 = cur[0][3], ref[0][3]
        eM = max(abs((c[3] - m0) / m0 - (q[3] - q0) / q0) for c, q in zip(cur, ref))
        mono = all(cur[i + 1][1] <= cur[i][1] * 1.0000001 for i in range(len(cur) - 1))
        out[f] = {"peak_tstar": cur[ip][0], "peak_offset_tstar": cur[ip][0] - ref[ipk][0],
                  "peak_enstrophy": cur[ip][2], "max_rel_energy_err": eE, "max_rel_enstrophy_err": eZ,
                  "max_mass_drift_diff": eM, "E_end_over_E0": cur[-1][1] / E0, "energy_monotonic": mono,
                  "finite": all(math.isfinite(c[1]) and math.isfinite(c[2]) for c in cur), "wall_s": r["wall_s"]}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--re", type=float, default=1600.0)
    ap.add_argument("--u0", type=float, default=0.05)
    ap.add_argument("--tstar", type=float, default=12.0)
    ap.add_argument("--every", type=int, default=20)
    ap.add_argument("--amr", type=int, default=0, help="also run the 3-level AMR with this coarse N0")
    ap.add_argument("--adapt", type=int, default=10)
    ap.add_argument("--out", default="precision_study.json")
    args = ap.parse_args()
    res = {"config": vars(args)}
    runs = {}
    for f in FORMS:
        print(f"# uniform {args.n}^3 {f}", flush=True)
        runs[f] = run_uniform(args.n, args.re, args.u0, args.tstar, args.every, f)
    res["uniform"] = {"summary": summarise(runs), "curves": {f: r["curve"] for f, r in runs.items()}}
    if args.amr:
        aruns = {}
        for f in FORMS:
            print(f"# amr N0={args.amr} {f}", flush=True)
            aruns[f] = run_amr(args.amr, args.re, args.u0, args.tstar, args.every, f, args.adapt,
                               None if f == "fp64" else aruns["fp64"]["script"])
        res["amr"] = {"summary": summarise(aruns), "curves": {f: r["curve"] for f, r in aruns.items()},
                      "counts": {f: r["counts"] for f, r in aruns.items()}}
    with open(args.out, "w") as fh:
        json.dump(res, fh)
    for part in ("uniform", "amr"):
        if part not in res:
            continue
        print(f"\n{part}: form | peak t* | d(peak) vs fp64 | max rel E err | max rel enstrophy err | mass drift diff")
        for f, s in res[part]["summary"].items():
            print(f"  {f:5s} | {s['peak_tstar']:.3f} | {s['peak_offset_tstar']:+.3f} | {s['max_rel_energy_err']:.2e} | "
                  f"{s['max_rel_enstrophy_err']:.2e} | {s['max_mass_drift_diff']:.2e}")


if __name__ == "__main__":
    main()
