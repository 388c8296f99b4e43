"""The interface pins catch plausible mistakes: each case plants one mistake in a copy of the oracle's
sources (on the CPU, in a temporary directory), builds it, runs tests/interface_pins.py against that build
in a child process, and checks that the pins the case names fail while the unplanted oracle passes them
all (tests/test_oracle_interface.py)."""
import concurrent.futures as cf
import json
import os
import shutil
import subprocess
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
PINS = os.path.join(ROOT, "tests", "interface_pins.py")

# (name, function anchor in oracle/amr.hpp, text after the anchor, planted text, pins that must fail)
MUTANTS = [
    ("prolong_L0 weights swapped", "void prolong_L0(", "(di ? tx : 1 - tx)", "(di ? 1 - tx : tx)",
     ["pin_seed_l1", "pin_migrate"]),
    ("prolong_L1 weights swapped", "void prolong_L1(", "(dj ? ty : 1 - ty)", "(dj ? 1 - ty : ty)",
     ["pin_seed_l2"]),
    ("prolong_prep_L0 weights swapped", "void prolong_prep_L0(", "(dk ? tz : 1 - tz)", "(dk ? 1 - tz : tz)",
     ["pin_fill_l1"]),
    ("prolong_prep_L1 weights swapped", "void prolong_prep_L1(", "(di ? tx : 1 - tx)", "(di ? 1 - tx : tx)",
     ["pin_fill_l2"]),
    ("level-1 ghost position off by one", "void fillcf_L1p()", "int gfi = pmod(ox + i, N1)",
     "int gfi = pmod(ox + i + 1, N1)", ["pin_fill_l1"]),
    ("level-2 ghost position off by one", "void fillcf_L2p()", "g2k = pmod(oz + k, N2)",
     "g2k = pmod(oz + k - 1, N2)", ["pin_fill_l2"]),
    ("level-1 step reads the mirrored ghost", "void step_L1()", "int ei = i - cvx[q] + 1",
     "int ei = i + cvx[q] + 1", ["pin_ghost_read_l1"]),
    ("level-2 step reads the mirrored ghost", "void step_L2()", "ek = k - cvz[q] + 1", "ek = k + cvz[q] + 1",
     ["pin_ghost_read_l2"]),
    ("prep without nqs", "void prep_list(", "+ R(nqs) * nq[q]", "+ nq[q]", ["pin_fill_l1", "pin_fill_l2"]),
    ("prolongation without nqs", "void prolong_L0(", "+ R(nqs) * nq[q]", "+ nq[q]", ["pin_seed_l1"]),
    ("restriction without nqi", "void restrict_level(", "+ R(nqi) * nq[q]", "+ nq[q]",
     ["pin_restrict_l1", "pin_restrict_l2"]),
    ("restriction octant transposed", "void restrict_level(", "oct = qz * 4 + qy * 2 + qx",
     "oct = qx * 4 + qy * 2 + qz", ["pin_restrict_l1", "pin_restrict_l2"]),
    ("restriction child offset", "void restrict_level(", "const int fi0 = 2 * (lci % HALF)",
     "const int fi0 = (2 * (lci % HALF) + 1) % MB", ["pin_restrict_l1", "pin_restrict_l2"]),
    ("sensor without 1/rho", "void sensor(", "float v = float(std::sqrt(s) / rho)", "float v = float(std::sqrt(s))",
     ["pin_sensor"]),
    ("sensor takes momentum for velocity", "void sensor(", "ux /= rho;", "", ["pin_sensor"]),
    ("absent corners not renormalised", "void prolong_prep_L1(", "if (wsum > 0.f && wsum < 0.999f)", "if (false)",
     ["pin_renormalised_fill"]),
    ("level-1 migration copies from the new rank", "void migrate_L1()", "int ofb = o_fof0[cb] + oct",
     "int ofb = fof0[cb] + oct", ["pin_migrate"]),
    ("level-2 migration keyed by the new level-1 index", "void migrate_L2()", "const int old_l1b = o_l1grid[p]",
     "const int old_l1b = l1grid[p]", ["pin_migrate"]),
]


def plant(src, anchor, old, new):
    """src with the first `old` inside the member function that starts at `anchor` replaced."""
    i = src.index(anchor)
    ends = [e for e in (src.find("\n  void ", i + 1), src.find("\n  static ", i + 1), src.find("\n  bool ", i + 1))
            if e > 0]
    j = src.index(old, i)
    assert j < min(ends, default=len(src)), f"{old!r} is not inside {anchor!r}"
    return src[:j] + new + src[j + len(old):]


def build_and_run(tmp, k, mutant):
    name, anchor, old, new, _ = mutant
    d = os.path.join(tmp, f"m{k}")
    os.makedirs(d)
    for f in ("capi.cpp", "oracle.hpp"):
        shutil.copy(os.path.join(ORACLE, f), d)
    with open(os.path.join(ORACLE, "amr.hpp")) as fh:
        src = fh.read()
    with open(os.path.join(d, "amr.hpp"), "w") as fh:
        fh.write(plant(src, anchor, old, new))
    so = os.path.join(d, "liboracle.so")
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared",
                    os.path.join(d, "capi.cpp"), "-o", so], check=True, capture_output=True)
    env = dict(os.environ, ORACLE_LIB=so, OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, PINS], env=env, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        return {"crash": r.stderr[-2000:]}
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.fixture(scope="module")
def outcomes():
    tmp = tempfile.mkdtemp(prefix="orc_mut_")
    try:
        with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
            futs = {m[0]: ex.submit(build_and_run, tmp, k, m) for k, m in enumerate(MUTANTS)}
            yield {n: f.result() for n, f in futs.items()}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def test_every_mutation_site_exists():
    with open(os.path.join(ORACLE, "amr.hpp")) as fh:
        src = fh.read()
    for name, anchor, old, new, _ in MUTANTS:
        assert plant(src, anchor, old, new) != src, name


@pytest.mark.parametrize("mutant", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_pins_catch_planted_mistake(outcomes, mutant):
    name, _, _, _, must_fail = mutant
    res = outcomes[name]
    assert "crash" not in res, res
    missed = [p for p in must_fail if res.get(p) is None]
    assert not missed, f"{name}: pins {missed} still pass"
