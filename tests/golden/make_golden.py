"""Writes golden.json: small known-answer fixtures of the oracle (the reference ships none).

Each case records SHA-256 digests of the stored population bytes (in the reference's tile-major layout)
and the diagnostics after the TGV initial state and after a few steps, plus the AMR block tables of a
sensor-free grid (--l1 slab --l2 all, whose tables do not depend on floating point). The digests freeze
the oracle's output: tests/test_golden.py checks the oracle against them on the CPU, and
tests/test_gpu_golden.py checks the device against them through the C ABI, both bit for bit.

Run from the repository root:  python3 tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as orc  # noqa: E402

PREC_NAMES = {orc.PREC_FP32: "fp32", orc.PREC_FP16: "fp16", orc.PREC_BF16: "bf16", orc.PREC_FP64: "fp64"}
# uniform cases: (edge, storage form, steps after init at which a record is taken)
UNIFORM = [(8, p, [1, 5]) for p in PREC_NAMES] + [(16, p, [1, 5]) for p in PREC_NAMES] + [(32, orc.PREC_FP32, [2, 5])]
# AMR case: coarse edge 32, --l1 slab --l2 all (src/tgv3d_amr.cu:777-797): half of level 0 refined, every
# deep level-1 tile refined again, so both coarse-fine interfaces exist; Re 800, 64 GiB budget
AMR = dict(n=32, re=800.0, u0=0.05, l1=orc.MODE_SLAB, l2=orc.MODE_ALL, frac1=0.35, frac2=0.35)
AMR_STEPS = [1, 2]
AMR_BUDGET = 64.0 * 2 ** 30
TABLES = ["ref0", "fof0", "rl0", "l1pcb", "l1oct", "f1pos", "l1grid", "ref1", "fof1", "rl1", "l2pl1", "l2oct",
          "skin0", "skin1"]
COUNTS = ["cnb0", "nref0", "nf1", "nref1", "nf2", "n_skin0", "n_skin1"]


def digest(a):
    return hashlib.sha256(a.tobytes()).hexdigest()


def uniform_case(n, prec, marks):
    o = orc.Uniform(n, re=1600.0, prec=prec)
    o.init()
    rec = {"n": n, "prec": PREC_NAMES[prec], "re": 1600.0, "u0": 0.05,
           "records": [{"step": 0, "sha256": digest(o.raw()), "diag": [float(v) for v in o.diag()]}]}
    done = 0
    for m in marks:
        o.step(m - done)
        done = m
        rec["records"].append({"step": m, "sha256": digest(o.raw()), "diag": [float(v) for v in o.diag()]})
    return rec


def amr_case():
    o = orc.Amr(budget_bytes=AMR_BUDGET, **AMR)
    o.init()
    c = o.counts()
    rec = {"params": {**AMR, "budget_gib": 64.0}, "counts": {k: int(c[k]) for k in COUNTS},
           "tables": {t: digest(o.table(t)) for t in TABLES}, "records": []}
    rec["records"].append({"step": 0, "sha256": [digest(o.raw(lvl)) for lvl in range(3)],
                           "diag": [float(v) for v in o.diag()]})
    done = 0
    for m in AMR_STEPS:
        o.step(m - done)
        done = m
        rec["records"].append({"step": m, "sha256": [digest(o.raw(lvl)) for lvl in range(3)],
                               "diag": [float(v) for v in o.diag()]})
    return rec


def main():
    out = {"note": "written by tests/golden/make_golden.py from the oracle; digests of stored population bytes "
                   "in the tile-major layout",
           "uniform": [uniform_case(n, p, m) for n, p, m in UNIFORM],
           "amr": amr_case()}
    path = os.path.join(HERE, "golden.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
