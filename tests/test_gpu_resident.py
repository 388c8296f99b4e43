"""Regrids on device-resident counts (graphs on, sensor-driven AMR), through the C ABI.

With graphs on, jp_regrid reads no count back: every level-1 and level-2 launch sizes itself from the
tile counts in device memory, the pools are sized for the largest counts a regrid can produce, and the
captured pairs of steps stay valid across regrids. The bar: the same run as the eager path (host-read
counts, recapture after every regrid) bit for bit, in tables, counts and populations, and the oracle's
tables and populations bit for bit; a count beyond a budget-limited pool is reported, never written.
"""
import os
import subprocess

import numpy as np
import pytest

import jpabi
import oracle as orc

pytestmark = pytest.mark.gpu
TABLES = ["ref0", "fof0", "rl0", "l1pcb", "l1oct", "f1pos", "l1grid", "ref1", "fof1", "rl1", "l2pl1", "l2oct",
          "skin0", "skin1", "sen0", "sen1"]
# the pool capacities differ by design (resident pools hold the largest possible counts)
COUNTS = ["cnb0", "nref0", "nf1", "nref1", "nf2", "n_skin0", "n_skin1", "cap_nref0", "cap_nref1", "hw_nf1", "hw_nf2",
          "ndeep"]
PB = 19 * 4 * (2 * 512 + 1000) * 6 / 5  # pool bytes per fine tile as the budget counts them


def dynamic(n0, prec=jpabi.PREC_FP32, **kw):
    args = dict(n=n0, re=1600.0, l1=jpabi.MODE_SENSOR, l2=jpabi.MODE_SENSOR, frac1=0.3, frac2=0.3, maxgb=16.0,
                prec=prec)
    args.update(kw)
    return jpabi.Amr(**args)


def same_grid(a, b, counts=COUNTS):
    ca, cb = a.counts(), b.counts()
    for k in counts:
        assert ca[k] == cb[k], (k, ca[k], cb[k])
    for t in TABLES:
        assert np.array_equal(a.table(t), b.table(t)), t


def same_populations(a, b):
    for lvl in range(3):
        assert np.array_equal(a.raw(lvl), b.raw(lvl)), f"level {lvl}"


def sensors(n0, seed):
    rng = np.random.default_rng(seed)
    return rng.random((n0 // 8) ** 3).astype(np.float32), rng.random((n0 // 4) ** 3).astype(np.float32)


@pytest.mark.parametrize("prec", [jpabi.PREC_FP32, jpabi.PREC_FP16])
def test_resident_regrids_match_eager_bit_exact(prec):
    n0 = 32
    s0, s1 = sensors(n0, 3)
    e, r = dynamic(n0, prec), dynamic(n0, prec)
    e.init(s0, s1)
    r.init(s0, s1)
    r.use_graph(True)  # pools resized for resident regrids, populations carried over
    cap = r.counts()
    assert cap["cap1"] == 8 * cap["cnb0"] and cap["cap2"] >= 8 * cap["cap1"]
    same_populations(e, r)
    for k in range(3):
        e.step(4)
        r.step(4)  # two replays of a captured pair
        e.regrid()
        r.regrid()  # device-resident: no count read, no recapture
        same_grid(e, r)
        same_populations(e, r)
    e.step(3)
    r.step(3)  # one pair and one eager step
    same_populations(e, r)


def test_resident_requested_before_init_and_odd_steps():
    n0 = 24
    s0, s1 = sensors(n0, 11)
    e, r = dynamic(n0), dynamic(n0)
    r.use_graph(True)  # applied by the initial build
    e.init(s0, s1)
    r.init(s0, s1)
    for n in (1, 2, 5):
        e.step(n)
        r.step(n)
        e.regrid()
        r.regrid()
    same_grid(e, r)
    same_populations(e, r)


def test_resident_run_follows_oracle():
    # as the eager dynamic-run test: the first marking fed the oracle's sensors, then the GPU computes its
    # own sensors in every regrid, on device-resident counts, and must reproduce the oracle's grid
    n0 = 24
    g = dynamic(n0)
    o = orc.Amr(n=n0, re=1600.0, l1=orc.MODE_SENSOR, l2=orc.MODE_SENSOR, frac1=0.3, frac2=0.3,
                budget_bytes=16.0 * 2 ** 30)
    assert o.init() == 0
    sen1pos = np.zeros((n0 // 4) ** 3, np.float32)
    sen1pos[o.table("f1pos")] = o.table("sen1")
    g.init(o.table("sen0"), sen1pos)
    for lvl in range(3):
        g.set_raw(o.raw(lvl), lvl)
    g.use_graph(True)
    for _ in range(3):
        o.step(4)
        g.step(4)
        g.regrid()
        assert o.regrid() == 0
        gc, oc = g.counts(), o.counts()
        for k in ["nref0", "nf1", "nref1", "nf2", "n_skin0", "n_skin1", "hw_nf1", "hw_nf2"]:
            assert gc[k] == oc[k], (k, gc[k], oc[k])
        for t in TABLES:
            assert np.array_equal(g.table(t), o.table(t)), t
        for lvl in range(3):
            assert np.array_equal(g.raw(lvl), o.raw(lvl)), f"level {lvl}"


def test_graph_bench_regrids_like_the_eager_bench():
    # the timed loop with graphs regrids on the same schedule (adapt 3: odd segments) on device-resident
    # counts; work is summed from counts logged per segment
    n0 = 24
    s0, s1 = sensors(n0, 7)
    e, r = dynamic(n0, adapt=3), dynamic(n0, adapt=3)
    e.init(s0, s1)
    r.init(s0, s1)
    re_ = e.bench(10, warmup=4, graph=False)
    rr = r.bench(10, warmup=4, graph=True)
    assert rr["steps"] == re_["steps"] == 10
    assert rr["work"] == re_["work"]
    same_grid(e, r)
    same_populations(e, r)


def test_resident_overflow_is_reported_not_written():
    # a budget too small for the largest level-1 count; a state at rest ties every level-0 tile at the
    # threshold, so the resident regrid's count exceeds the pool: the overflowing parents stay unrefined
    # (steps stay in bounds) and the next count read reports the overflow
    n0 = 24
    cnb0 = (n0 // 8) ** 3
    s0, s1 = sensors(n0, 5)
    g = dynamic(n0, frac1=0.1, frac2=0.1, maxgb=120 * PB / 2 ** 30)
    g.init(s0, s1)
    g.use_graph(True)
    c = g.counts()
    assert c["cap1"] < 8 * cnb0
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, dtype=np.float32)
    for lvl in range(3):
        nt = g.raw(lvl).size // (19 * 512)
        g.set_raw(np.broadcast_to(w[None, :, None], (nt, 19, 512)).ravel(), lvl)
    g.regrid()
    g.step(2)
    with pytest.raises(jpabi.JPError) as err:
        g.counts()
    assert err.value.code == jpabi.JP_ERR_CAPACITY
    with pytest.raises(jpabi.JPError) as err:
        g.step(2)
    assert err.value.code == jpabi.JP_ERR_STATE


def test_amr_driver_graph_loop_matches_eager_loop():
    # the production loop of ./amr --graph (src/tgv3d_amr.cu:872-881): pairs of coarse steps from graphs,
    # regrids on device-resident counts between them; the diagnostics must equal the eager loop's
    build = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "jp-solver_amd", "build")
    args = [os.path.join(build, "amr"), "--n", "24", "--re", "1600", "--l1", "sensor", "--l2", "sensor", "--frac1", "0.3",
            "--frac2", "0.3", "--adapt", "4", "--tstar", "0.21", "--diag-every", "2", "--maxgb", "4"]
    out = []
    for graph in ("0", "1"):
        r = subprocess.run(args + ["--graph", graph], capture_output=True, text=True, timeout=120)
        assert r.returncode in (0, 1), r.stdout + r.stderr
        lines = [x for x in r.stdout.splitlines() if x.startswith("  t*=") or "enstrophy peak" in x]
        out.append(lines)
    assert len(out[0]) == 10 and out[0] == out[1]


def rest_state(ctx, set_raw):
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, dtype=np.float32)
    for lvl in range(3):
        nt = ctx.raw(lvl).size // (19 * 512)
        set_raw(lvl, np.broadcast_to(w[None, :, None], (nt, 19, 512)).ravel())


def test_eager_regrid_is_redone_when_ties_exceed_a_pool():
    # the eager regrid runs on device counts with one count read at the end; a state at rest ties every
    # tile at both thresholds, the counts exceed the pools, and the regrid is undone and run again with the
    # host growing the pools: the oracle (same growth rule) must agree bit for bit
    n0 = 24
    cnb0 = (n0 // 8) ** 3
    s0, s1 = sensors(n0, 5)
    g = dynamic(n0, frac1=0.1, frac2=0.1)
    o = orc.Amr(n=n0, re=1600.0, l1=orc.MODE_SENSOR, l2=orc.MODE_SENSOR, frac1=0.1, frac2=0.1,
                budget_bytes=16.0 * 2 ** 30)
    g.init(s0, s1)
    assert o.init(s0, s1) == 0
    cap1 = g.counts()["cap1"]
    rest_state(g, lambda lvl, a: g.set_raw(a, lvl))
    rest_state(o, lambda lvl, a: o.set_raw(lvl, a))
    g.regrid()
    assert o.regrid() == 0
    c = g.counts()
    assert c["nf1"] == 8 * cnb0 > cap1 and c["cap1"] >= c["nf1"]
    oc = o.counts()
    for k in ["nref0", "nf1", "nref1", "nf2", "cap1", "cap2", "hw_nf1", "hw_nf2"]:
        assert c[k] == oc[k], (k, c[k], oc[k])
    for t in TABLES:
        assert np.array_equal(g.table(t), o.table(t)), t
    g.step(2)
    o.step(2)
    for lvl in range(3):
        assert np.array_equal(g.raw(lvl), o.raw(lvl)), f"level {lvl}"
