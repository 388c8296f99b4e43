"""The oracle against its committed known-answer fixtures (tests/golden/golden.json, written by
tests/golden/make_golden.py). The reference ships no golden vectors, so these freeze the oracle's own
output: SHA-256 digests of the stored population bytes after the TGV initial state and a few steps, in
every storage form, and the block tables, counts and populations of a 3-level grid (coarse 32, --l1 slab
--l2 all). The oracle itself is pinned by the closed forms and invariants of the other oracle tests; these
digests make any later change to its output visible, and the device is checked against the same digests
in test_gpu_golden.py."""
import hashlib
import json
import math
import os

import pytest

import oracle as orc

GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.json")))
PREC = {"fp32": orc.PREC_FP32, "fp16": orc.PREC_FP16, "bf16": orc.PREC_BF16, "fp64": orc.PREC_FP64}


def digest(a):
    return hashlib.sha256(a.tobytes()).hexdigest()


@pytest.mark.parametrize("case", GOLD["uniform"], ids=lambda c: f"{c['n']}-{c['prec']}")
def test_oracle_uniform_matches_fixture(case):
    o = orc.Uniform(case["n"], re=case["re"], u0=case["u0"], prec=PREC[case["prec"]])
    o.init()
    done = 0
    for r in case["records"]:
        if r["step"] > done:
            o.step(r["step"] - done)
            done = r["step"]
        assert digest(o.raw()) == r["sha256"], r["step"]
        assert list(o.diag()) == pytest.approx(r["diag"], rel=1e-13)


def test_fixture_initial_diagnostics_are_the_closed_form():
    """The fixtures start where the mathematics says: E0 = u0^2/8, enstrophy0 = 3/8 u0^2 sin^2(2 pi/N) and
    mass 1, to the storage form's rounding (16-bit forms store a scaled deviation from the weight)."""
    tol = {"fp64": 1e-12, "fp32": 1e-6, "fp16": 1e-3, "bf16": 2e-3}
    mtol = {"fp64": 1e-12, "fp32": 1e-7, "fp16": 2e-6, "bf16": 1e-5}
    for case in GOLD["uniform"]:
        E0, ens0, mass = case["records"][0]["diag"]
        u0, n = case["u0"], case["n"]
        assert E0 == pytest.approx(u0 ** 2 / 8, rel=tol[case["prec"]])
        assert ens0 == pytest.approx(3 / 8 * u0 ** 2 * math.sin(2 * math.pi / n) ** 2, rel=tol[case["prec"]])
        assert mass == pytest.approx(1.0, abs=mtol[case["prec"]])


def test_oracle_amr_matches_fixture():
    a = GOLD["amr"]
    p = dict(a["params"])
    budget = p.pop("budget_gib") * 2 ** 30
    o = orc.Amr(budget_bytes=budget, **p)
    o.init()
    c = o.counts()
    for k, v in a["counts"].items():
        assert c[k] == v, k
    for t, h in a["tables"].items():
        assert digest(o.table(t)) == h, t
    done = 0
    for r in a["records"]:
        if r["step"] > done:
            o.step(r["step"] - done)
            done = r["step"]
        assert [digest(o.raw(lvl)) for lvl in range(3)] == r["sha256"], r["step"]
        assert list(o.diag()) == pytest.approx(r["diag"], rel=1e-13)
