"""The device against the committed known-answer fixtures (tests/golden/golden.json), through the C ABI.

Bar: bit for bit. Each uniform case starts from the oracle's TGV initial state (whose digest the fixture
holds; the device's own initial state may differ from it in the last bit of a few values, from sin/cos),
then steps on the device; the stored populations' SHA-256 digest must equal the fixture's after every
record. 32^3 FP32 runs as fused step pairs, the other edges as single steps. The AMR case (coarse 32,
--l1 slab --l2 all) builds its grid on the device: counts and table digests must equal the fixture's;
from the oracle's initial populations, the coarse steps must reproduce the fixture's digests on all three
levels."""
import hashlib
import json
import os

import pytest

import jpabi
import oracle as orc

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.json")))
PREC = {"fp32": 0, "fp16": 1, "bf16": 2, "fp64": 3}


def digest(a):
    return hashlib.sha256(a.tobytes()).hexdigest()


@pytest.mark.parametrize("case", GOLD["uniform"], ids=lambda c: f"{c['n']}-{c['prec']}")
def test_device_uniform_matches_fixture(case):
    prec = PREC[case["prec"]]
    o = orc.Uniform(case["n"], re=case["re"], u0=case["u0"], prec=prec)
    o.init()
    start = o.raw()
    assert digest(start) == case["records"][0]["sha256"]
    g = jpabi.Uniform(case["n"], re=case["re"], u0=case["u0"], prec=prec)
    g.set_raw(start)
    done = 0
    for r in case["records"][1:]:
        g.step(r["step"] - done)
        done = r["step"]
        assert digest(g.raw()) == r["sha256"], r["step"]
        assert list(g.diag()) == pytest.approx(r["diag"], rel=1e-12)
    g.close()


def test_device_amr_matches_fixture():
    a = GOLD["amr"]
    p = dict(a["params"])
    gib = p.pop("budget_gib")
    g = jpabi.Amr(maxgb=gib, **p)
    g.init()
    c = g.counts()
    for k, v in a["counts"].items():
        assert c[k] == v, k
    for t, h in a["tables"].items():
        assert digest(g.table(t)) == h, t
    o = orc.Amr(budget_bytes=gib * 2 ** 30, **p)
    o.init()
    for lvl in range(3):
        g.set_raw(o.raw(lvl), lvl)
    assert [digest(g.raw(lvl)) for lvl in range(3)] == a["records"][0]["sha256"]
    done = 0
    for r in a["records"][1:]:
        g.step(r["step"] - done)
        done = r["step"]
        assert [digest(g.raw(lvl)) for lvl in range(3)] == r["sha256"], r["step"]
        assert list(g.diag()) == pytest.approx(r["diag"], rel=1e-12)
    g.close()
