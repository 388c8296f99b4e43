"""GPU parity of the 3-level AMR against the CPU oracle, through the C ABI.

Bar: block tables, marks, scans, skins and the counts are bit-exact (given the same sensor field, which
is injected through the test hook where the sensor is not the object of the test); from the same stored
state on every level, the coarse step of the 1:2:4 subcycle (prep, ghost fill, fine steps, restriction)
is bit-exact in every storage form; the computed sensors are bit-exact; runs from each side's own
initial state agree within 1e-5 relative; the composite diagnostics within 1e-12.
"""
import numpy as np
import pytest

import jpabi
import oracle as orc

pytestmark = pytest.mark.gpu
PRECS = [jpabi.PREC_FP32, jpabi.PREC_FP16, jpabi.PREC_BF16, jpabi.PREC_FP64]
TABLES = ["ref0", "fof0", "rl0", "l1pcb", "l1oct", "f1pos", "l1grid", "ref1", "fof1", "rl1", "l2pl1", "l2oct",
          "skin0", "skin1"]
COUNTS = ["cnb0", "nref0", "nf1", "nref1", "nf2", "n_skin0", "n_skin1", "cap_nref0", "cap1", "cap_nref1", "cap2",
          "hw_nf1", "hw_nf2"]


def pair(**kw):
    okw = dict(kw)
    okw.pop("adapt", None)
    maxgb = okw.pop("maxgb", 0.0)
    if maxgb:
        okw["budget_bytes"] = maxgb * 2 ** 30
    else:
        kw["maxgb"] = 64.0
        okw["budget_bytes"] = 64.0 * 2 ** 30
    return jpabi.Amr(**kw), orc.Amr(**okw)


def same_tables(g, o):
    gc, oc = g.counts(), o.counts()
    for k in COUNTS:
        assert gc[k] == oc[k], (k, gc[k], oc[k])
    for t in TABLES:
        a, b = g.table(t), o.table(t)
        assert np.array_equal(a, b), t


def copy_state(o, g):
    for lvl in range(3):
        g.set_raw(o.raw(lvl), lvl)


def same_state(g, o, exact=True, rtol=1e-5):
    for lvl in range(3):
        a, b = g.raw(lvl), o.raw(lvl)
        assert a.size == b.size, lvl
        if exact:
            assert np.array_equal(a, b), f"level {lvl}"
        else:
            a, b = g_f64(g, lvl, o), o.f64(lvl)
            assert np.max(np.abs(a - b) / np.abs(b)) <= rtol, f"level {lvl}"


def g_f64(g, lvl, o):
    return decode(g.raw(lvl), g.prec, o.scales[lvl])


def decode(raw, prec, scale):
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, dtype=np.float32)
    nt = raw.size // (19 * 512)
    if prec in (jpabi.PREC_FP32, jpabi.PREC_FP64):
        return raw.astype(np.float64)
    if prec == jpabi.PREC_FP16:
        v = raw.view(np.float16).astype(np.float32)
    else:
        v = (raw.astype(np.uint32) << 16).view(np.float32)
    v = v.reshape(nt, 19, 512) * np.float32(1.0 / scale) + w[None, :, None]
    return v.ravel().astype(np.float64)


@pytest.mark.parametrize("l1,l2", [(jpabi.MODE_ALL, jpabi.MODE_ALL), (jpabi.MODE_SLAB, jpabi.MODE_ALL)])
def test_initial_tables_match(l1, l2):
    g, o = pair(n=32, re=800.0, l1=l1, l2=l2)
    g.init()
    assert o.init() == 0
    same_tables(g, o)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("l1,l2", [(jpabi.MODE_ALL, jpabi.MODE_ALL), (jpabi.MODE_SLAB, jpabi.MODE_ALL)])
def test_subcycle_is_bit_exact(prec, l1, l2):
    n0 = 16 if l1 == jpabi.MODE_ALL else 32
    g, o = pair(n=n0, re=800.0, l1=l1, l2=l2, prec=prec)
    g.init()
    o.init()
    copy_state(o, g)
    for _ in range(2):
        o.step(1)
        g.step(1)
        same_state(g, o)


@pytest.mark.parametrize("prec", PRECS)
def test_random_states_through_the_interface_are_bit_exact(prec):
    """Every level set to populations drawn independently per cell and direction (FP32 / FP64 f = w_q u,
    u uniform in [0.2, 3]: densities 0.5 - 2.5, arbitrary momenta; FP16 / BF16 stored deviations uniform in
    +-600), then two coarse steps on a grid with both interfaces (coarse 32, --l1 slab --l2 all): the fills'
    prep, the fine steps and the restrictions, whose velocity quotients come from one reciprocal, agree
    with the oracle's IEEE division bit for bit."""
    g, o = pair(n=32, re=800.0, l1=jpabi.MODE_SLAB, l2=jpabi.MODE_ALL, prec=prec)
    g.init()
    o.init()
    rng = np.random.default_rng(11 + prec)
    for lvl in range(3):
        size = o.raw(lvl).size
        q = (np.arange(size) // 512) % 19
        w = np.where(q == 0, 1 / 3, np.where(q < 7, 1 / 18, 1 / 36))
        if prec == jpabi.PREC_FP64:
            a = w * rng.uniform(0.2, 3.0, size)
        elif prec == jpabi.PREC_FP32:
            a = (w * rng.uniform(0.2, 3.0, size)).astype(np.float32)
        else:
            d = rng.uniform(-600.0, 600.0, size)
            a = d.astype(np.float16).view(np.uint16) if prec == jpabi.PREC_FP16 else \
                (d.astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)
        o.set_raw(lvl, a)
        g.set_raw(a, lvl)
    for _ in range(2):
        o.step(1)
        g.step(1)
        same_state(g, o)


def test_own_init_run_within_tolerance():
    g, o = pair(n=32, re=800.0, l1=jpabi.MODE_SLAB, l2=jpabi.MODE_ALL)
    g.init()
    o.init()
    same_state(g, o, exact=False, rtol=1e-5)
    g.step(4)
    o.step(4)
    same_state(g, o, exact=False, rtol=1e-5)
    assert g.diag() == pytest.approx(o.diag(), rel=1e-9)


@pytest.mark.parametrize("prec", [jpabi.PREC_FP32, jpabi.PREC_FP64])
def test_composite_diag_matches(prec):
    g, o = pair(n=16, re=800.0, l1=jpabi.MODE_SLAB, l2=jpabi.MODE_ALL, prec=prec)
    g.init()
    o.init()
    o.step(2)
    copy_state(o, g)
    assert g.diag() == pytest.approx(o.diag(), rel=1e-12)


def test_sensor_mode_init_tables_bit_exact_with_injected_sensor():
    n0 = 32
    cnb0, ng1 = (n0 // 8) ** 3, (n0 // 4) ** 3
    rng = np.random.default_rng(21)
    sen0 = rng.integers(0, 40, cnb0).astype(np.float32)  # with ties
    sen1 = rng.integers(0, 40, ng1).astype(np.float32)
    g, o = pair(n=n0, re=1600.0, l1=jpabi.MODE_SENSOR, l2=jpabi.MODE_SENSOR, frac1=0.3, frac2=0.3)
    g.init(sen0, sen1)
    assert o.init(sen0, sen1) == 0
    same_tables(g, o)
    assert g.counts()["nf2"] > 0


def test_computed_sensor_is_bit_exact():
    n0 = 32
    g, o = pair(n=n0, re=1600.0, l1=jpabi.MODE_SLAB, l2=jpabi.MODE_ALL)
    g.init()
    o.init()
    o.step(3)
    copy_state(o, g)
    # the sensor values of the regrid are recorded in the sen0 / sen1 tables
    g.regrid()
    assert o.regrid() == 0
    assert np.array_equal(g.table("sen0"), o.table("sen0"))
    assert np.array_equal(g.table("sen1"), o.table("sen1"))
    same_tables(g, o)
    same_state(g, o)


def test_regrid_with_injected_sensors_migrates_bit_exact():
    n0 = 32
    cnb0, ng1 = (n0 // 8) ** 3, (n0 // 4) ** 3
    rng = np.random.default_rng(22)
    g, o = pair(n=n0, re=1600.0, l1=jpabi.MODE_SENSOR, l2=jpabi.MODE_SENSOR, frac1=0.4, frac2=0.3)
    s0, s1 = rng.random(cnb0).astype(np.float32), rng.random(ng1).astype(np.float32)
    g.init(s0, s1)
    o.init(s0, s1)
    copy_state(o, g)
    for it in range(3):
        o.step(2)
        g.step(2)
        same_state(g, o)
        s0 = rng.random(cnb0).astype(np.float32)
        s1 = rng.random(ng1).astype(np.float32)
        g.regrid(s0, s1)
        assert o.regrid(s0, s1) == 0
        same_tables(g, o)
        same_state(g, o)


def test_dynamic_sensor_run_tables_follow_oracle():
    # sensor mode end to end: develop(300), computed sensors, regrids. The GPU's first marking is fed the
    # oracle's computed sensors (the develop(300) states differ by the initial sin/cos ulps); from the
    # shared state on, the GPU computes its own sensors and must reproduce the oracle's tables.
    n0 = 24
    g, o = pair(n=n0, re=1600.0, l1=jpabi.MODE_SENSOR, l2=jpabi.MODE_SENSOR, frac1=0.3, frac2=0.3)
    assert o.init() == 0
    sen1pos = np.zeros((n0 // 4) ** 3, np.float32)
    sen1pos[o.table("f1pos")] = o.table("sen1")
    g.init(o.table("sen0"), sen1pos)
    same_tables(g, o)
    copy_state(o, g)
    for _ in range(2):
        o.step(5)
        g.step(5)
        g.regrid()
        assert o.regrid() == 0
        assert np.array_equal(g.table("sen0"), o.table("sen0"))
        same_tables(g, o)
        same_state(g, o)


PB = 19 * 4 * (2 * 512 + 1000) * 6 / 5  # pool bytes per fine tile as the budget counts them


def test_capacity_overflow_is_an_error_not_a_write():
    n0 = 24
    cnb0 = (n0 // 8) ** 3
    cap_nref0 = int(np.ceil(0.1 * cnb0 * 1.4))
    # the budget holds only the target-sized level-1 pool; every tile tied at the threshold exceeds it
    g = jpabi.Amr(n=n0, re=1600.0, l1=jpabi.MODE_SENSOR, l2=jpabi.MODE_SENSOR, frac1=0.1, frac2=0.1,
                  maxgb=8 * cap_nref0 * PB * 1.0001 / 2 ** 30)
    with pytest.raises(jpabi.JPError) as e:
        g.init(np.ones(cnb0, np.float32))
    assert e.value.code == jpabi.JP_ERR_CAPACITY


def test_level1_pool_grows_on_tied_init_and_regrid():
    """Threshold ties beyond the level-1 pool grow it (the reference would write out of bounds), at init
    and during a regrid whose migration still reads the old pool; bit-exact with the oracle's rule."""
    n0 = 32
    cnb0, ng1 = (n0 // 8) ** 3, (n0 // 4) ** 3
    rng = np.random.default_rng(5)
    s0, s1 = rng.random(cnb0).astype(np.float32), rng.random(ng1).astype(np.float32)
    for prec in (jpabi.PREC_FP32, jpabi.PREC_FP16):
        g, o = pair(n=n0, re=1600.0, prec=prec, l1=jpabi.MODE_SENSOR, l2=jpabi.MODE_SENSOR, frac1=0.1,
                    frac2=0.3)
        g.init(s0, s1)
        assert o.init(s0, s1) == 0
        cap1 = g.counts()["cap1"]
        copy_state(o, g)
        g.step(2)
        o.step(2)
        g.regrid(np.ones(cnb0, np.float32), s1)  # every coarse tile tied
        assert o.regrid(np.ones(cnb0, np.float32), s1) == 0
        c = g.counts()
        assert c["nf1"] == 8 * cnb0 > cap1 and c["cap1"] >= c["nf1"]
        same_tables(g, o)
        same_state(g, o)
        g.step(2)
        o.step(2)
        same_state(g, o)
    g, o = pair(n=24, re=1600.0, l1=jpabi.MODE_SENSOR, l2=jpabi.MODE_SENSOR, frac1=0.1, frac2=0.1)
    g.init(np.ones(27, np.float32))
    assert o.init(np.ones(27, np.float32)) == 0
    same_tables(g, o)


def test_amr_bench_and_graph_replay():
    g, o = pair(n=16, re=800.0, l1=jpabi.MODE_SLAB, l2=jpabi.MODE_ALL)
    g.init()
    o.init()
    copy_state(o, g)
    r = g.bench(4, warmup=0, graph=True)
    o.step(4)
    same_state(g, o)
    c = g.counts()
    assert r["work"] == pytest.approx(4 * (c["cnb0"] + 2 * c["nf1"] + 4 * c["nf2"]) * 512)


def test_level2_pool_grows_on_tied_regrid():
    """A regrid whose threshold ties exceed the level-2 pool grows the pool (the reference would write out
    of bounds); tables and populations stay bit-exact with the oracle, which applies the same rule."""
    n0 = 32
    cnb0, ng1 = (n0 // 8) ** 3, (n0 // 4) ** 3
    rng = np.random.default_rng(23)
    s0 = rng.random(cnb0).astype(np.float32)
    s1 = rng.random(ng1).astype(np.float32)
    g, o = pair(n=n0, re=1600.0, l1=jpabi.MODE_SENSOR, l2=jpabi.MODE_SENSOR, frac1=0.5, frac2=0.05)
    g.init(s0, s1)
    o.init(s0, s1)
    cap2 = g.counts()["cap2"]
    copy_state(o, g)
    g.step(2)
    o.step(2)
    s1 = np.ones(ng1, np.float32)  # every deep level-1 tile tied at the threshold
    g.regrid(s0, s1)
    assert o.regrid(s0, s1) == 0
    c = g.counts()
    assert c["nf2"] > cap2 and c["cap2"] >= c["nf2"]
    same_tables(g, o)
    same_state(g, o)
    g.step(2)
    o.step(2)
    same_state(g, o)
