"""The benchmark configurations run at their own sizes on the GPU, each against the bar it is quoted with:

* dynamic AMR, coarse 96 / finest 384, Re 1600, --l1/--l2 sensor, frac 0.3 / 0.1, regrid every 10 coarse
  steps, to t* = 12 through the ./amr driver: the validation gate of docs/validation.md:14 (breakdown
  present, energy monotonic, mass held); and a 20-coarse-step segment spanning two regrids, bit-exact
  against the oracle from a shared state (tables, counts, populations of every level);
* the precision form: FP16 and BF16 storage against an FP64 run of the same 128^3 lattice to t* = 12
  (FP16: enstrophy-peak time within 0.05 t*, energy within 0.5 %); the 768^3 lattice in FP16 and BF16 at
  full size (closed-form initial diagnostics, mass held, energy decaying, agreement with FP32); the
  dynamic AMR coarse 128 / finest 512 in FP16 and BF16 through steps and a regrid (block-table invariants,
  finite diagnostics, mass held);
* the 32-member seeded ensemble (jp-solver_amd/ensemble32.json): every member, after graph-replayed
  batched steps and a regrid, bit-identical to the same member advanced alone.
"""
import json
import math
import os
import subprocess

import numpy as np
import pytest

import jpabi
import oracle as orc
import precision_study as ps
from test_gpu_amr import same_state, same_tables
from test_gpu_ensemble import levels, standalone
from test_oracle_amr import check_table_invariants

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "jp-solver_amd", "build")


def sensors_by_pos(a, n0):
    s1 = np.zeros((n0 // 4) ** 3, np.float32)
    pos = a.table("f1pos")
    if pos.size:
        s1[pos] = a.table("sen1")
    return a.table("sen0"), s1


def test_dynamic_96_384_validation_gate():
    """docs/validation.md:14: the sensor-driven run shows the breakdown (enstrophy peak inside (4, 10)),
    decays monotonically and holds its mass (src/tgv3d_amr.cu:901)."""
    r = subprocess.run([os.path.join(BUILD, "amr"), "--n", "96", "--re", "1600", "--l1", "sensor", "--l2", "sensor",
                        "--frac1", "0.3", "--frac2", "0.1", "--adapt", "10", "--tstar", "12"],
                       capture_output=True, text=True, timeout=300)
    out = r.stdout
    assert r.returncode == 0, out[-3000:] + r.stderr
    assert "VALID (breakdown present, energy monotonic, mass held)" in out
    assert "steps=3667" in out.splitlines()[0]
    peak = float(out.split("enstrophy peak at t*=")[1].split()[0])
    assert 4.0 < peak < 10.0


def test_dynamic_96_segment_bit_exact_vs_oracle():
    """Coarse 96, 20 coarse steps with regrids at 10 and 20 (sensors computed by each side from the same
    state), from the device's developed initial state: every table, count and level bit-exact."""
    kw = dict(n=96, re=1600.0, l1=jpabi.MODE_SENSOR, l2=jpabi.MODE_SENSOR, frac1=0.3, frac2=0.1)
    g = jpabi.Amr(maxgb=64.0, adapt=10, **kw)
    g.init()  # develop(300) + the initial build on the device
    s0, s1 = sensors_by_pos(g, 96)
    o = orc.Amr(budget_bytes=64.0 * 2 ** 30, develop=0, **kw)
    assert o.init(s0, s1) == 0
    same_tables(g, o)
    for lvl in range(3):
        o.set_raw(lvl, g.raw(lvl))
    for s in range(1, 21):
        g.step(1)
        o.step(1)
        if s % 10 == 0:
            g.regrid()
            assert o.regrid() == 0
            same_tables(g, o)
    assert np.array_equal(g.table("sen0"), o.table("sen0"))
    assert np.array_equal(g.table("sen1"), o.table("sen1"))
    same_state(g, o)
    assert np.allclose(g.diag(), o.diag(), rtol=1e-12, atol=0)
    g.close()


@pytest.fixture(scope="module")
def precision_runs():
    return {f: ps.run_uniform(128, 1600.0, 0.05, 12.0, 20, f) for f in ("fp64", "fp32", "fp16", "bf16")}


def test_precision_forms_against_fp64_128(precision_runs):
    s = ps.summarise(precision_runs)
    print(json.dumps(s, indent=1))
    # (energy is not strictly monotonic at 128^3 in any form, FP64 included: Re 1600 is under-resolved on
    # this lattice; the reference's gate is stated for 192^3, tests/test_gpu_validation.py)
    for f in s:
        assert s[f]["finite"], f
    # the accuracy bar of the FP16 store: enstrophy-peak time within 0.05 t*, energy within 0.5 % of FP64
    assert abs(s["fp16"]["peak_offset_tstar"]) <= 0.05
    assert s["fp16"]["max_rel_energy_err"] <= 0.005
    assert s["fp16"]["max_mass_drift_diff"] <= 1e-5
    # FP32 keeps the reference's arithmetic; BF16 (8-bit significand) is reported, with a looser bound
    assert abs(s["fp32"]["peak_offset_tstar"]) <= 0.05 and s["fp32"]["max_rel_energy_err"] <= 1e-3
    assert abs(s["bf16"]["peak_offset_tstar"]) <= 0.1 and s["bf16"]["max_rel_energy_err"] <= 0.05
    assert 7.5 <= s["fp64"]["peak_tstar"] <= 10.0


# relative tolerances of energy, initial enstrophy, and enstrophy after 20 steps against FP32: the storage
# rounding biases the central-difference enstrophy by a term growing as N^2 (oracle, BF16: 5e-5 at 96^3, 2.1e-4
# at 192^3, 4.1e-4 at 256^3; FP16 about 100 x less), and each BF16 collision adds rounding noise at the grid
# scale, where the 768^3 lattice's own gradients are small (measured: 2.7 % after 20 steps)
TOL768 = {"fp16": (1e-4, 1e-3, 1e-3), "bf16": (1e-3, 1.5e-2, 6e-2)}


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_uniform_768_full_size(prec):
    n, u0 = 768, 0.05
    tol_e, tol_z, tol_z20 = TOL768[prec]
    g = jpabi.Uniform(n, re=1600.0, u0=u0, prec=jpabi.PREC_NAMES[prec])
    g.init()
    E0, ens0, m0 = g.diag()
    kf = 2 * math.pi / n
    # closed forms of the TGV initial state (mean of 1/2 |u|^2 = u0^2/8; central-difference enstrophy
    # 3/8 u0^2 sin^2 kf; mean density 1), to the storage's rounding
    assert E0 == pytest.approx(u0 ** 2 / 8, rel=tol_e)
    assert ens0 == pytest.approx(3 / 8 * u0 ** 2 * math.sin(kf) ** 2, rel=tol_z)
    assert m0 == pytest.approx(1.0, abs=1e-6)
    g.step(20)
    E1, ens1, m1 = g.diag()
    assert math.isfinite(E1) and E1 < E0
    assert abs(m1 - m0) / m0 < 1e-6  # the shifted store's deviation collision conserves mass
    g.close()
    f = jpabi.Uniform(n, re=1600.0, u0=u0, prec=jpabi.PREC_FP32)
    f.init()
    f.step(20)
    Ef, ensf, _ = f.diag()
    f.close()
    assert E1 == pytest.approx(Ef, rel=tol_e) and ens1 == pytest.approx(ensf, rel=tol_z20)


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_amr_128_512_steps_and_regrid(prec):
    a = jpabi.Amr(128, re=1600.0, u0=0.05, l1=jpabi.MODE_SENSOR, l2=jpabi.MODE_SENSOR, frac1=0.3, frac2=0.1, adapt=10,
                  prec=jpabi.PREC_NAMES[prec])
    a.init()
    c = a.counts()
    assert c["nf1"] > 0 and c["nf2"] > 0
    check_table_invariants(a, 128)
    d0 = a.diag()
    a.step(10)
    a.regrid()
    check_table_invariants(a, 128)
    a.step(2)
    d1 = a.diag()
    assert np.all(np.isfinite(d1)) and d1[0] < d0[0]
    assert abs(d1[2] - d0[2]) / d0[2] < 1e-4  # the coarse-fine interface is not exactly conservative
    a.close()


def test_ensemble32_members_match_alone():
    with open(os.path.join(ROOT, "jp-solver_amd", "ensemble32.json")) as fh:
        members = json.load(fh)["members"]
    assert len(members) == 32
    prec = jpabi.PREC_FP32
    ens = jpabi.Ensemble(members, prec=prec, maxgb=48.0)
    ens.init()
    alone = [standalone(m, prec) for m in members]
    for x in alone:
        x.init()
    ens.use_graph(True)

    def same():
        for k, (m, e, x) in enumerate(zip(members, ens.members, alone)):
            for lvl in range(levels(m)):
                assert np.array_equal(e.raw(lvl), x.raw(lvl)), (k, lvl)

    ens.step(4)
    for x in alone:
        x.step(4)
    same()
    ens.regrid()
    for m, x in zip(members, alone):
        if m["amr"]:
            x.regrid()
    ens.step(2)
    for x in alone:
        x.step(2)
    same()
    for x in alone:
        x.close()
    ens.close()
