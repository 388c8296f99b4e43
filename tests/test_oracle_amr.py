"""CPU pins of the oracle's 3-level AMR (src/tgv3d_amr.cu): level parameters in closed form, block-table
invariants checked by brute force, the deterministic slab fixture of the reference's marking rules, the
exact order-statistic threshold with ties, the rest state as a fixed point of the whole subcycle, and
migration as a bit copy for persisting tiles."""
import math

import numpy as np
import pytest

import oracle as orc

W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def pos3(p, g):
    return p % g, (p // g) % g, p // (g * g)


def test_level_parameters_static_gold():
    # tau doubles the non-equilibrium scaling per level: tau2 at N0 equals tau of the uniform 4 N0 lattice
    a = orc.Amr(64, re=800)
    p = a.params()
    nu4 = 0.05 * 256 / (2 * math.pi * 800)
    assert p["tau2"] == pytest.approx(3 * nu4 + 0.5, rel=1e-14)
    assert p["tau2"] == pytest.approx(0.507639, abs=1e-6)
    assert p["nqs01"] * p["nqi10"] == pytest.approx(1.0, rel=1e-6)
    assert p["nqs12"] * p["nqi21"] == pytest.approx(1.0, rel=1e-6)


def check_table_invariants(a, n0):
    c = a.counts()
    cgx, l1gx = n0 // 8, n0 // 4
    ref0, fof0, rl0 = a.table("ref0"), a.table("fof0"), a.table("rl0")
    l1pcb, l1oct, f1pos, l1grid = a.table("l1pcb"), a.table("l1oct"), a.table("f1pos"), a.table("l1grid")
    assert c["nref0"] == ref0.sum() and c["nf1"] == 8 * c["nref0"]
    assert np.array_equal(rl0, np.flatnonzero(ref0))  # ascending refined coarse tiles
    for r, cb in enumerate(rl0):
        assert fof0[cb] == 8 * r
    assert np.all(fof0[ref0 == 0] == -1)
    for b in range(c["nf1"]):
        cb, oct = l1pcb[b], l1oct[b]
        bx, by, bz = pos3(cb, cgx)
        assert f1pos[b] == ((2 * bz + (oct >> 2)) * l1gx + 2 * by + ((oct >> 1) & 1)) * l1gx + 2 * bx + (oct & 1)
        assert l1grid[f1pos[b]] == b
    assert (l1grid >= 0).sum() == c["nf1"]
    if c["nf1"]:
        ref1, fof1, rl1 = a.table("ref1"), a.table("fof1"), a.table("rl1")
        assert np.array_equal(rl1, np.flatnonzero(ref1))
        l2pl1, l2oct = a.table("l2pl1"), a.table("l2oct")
        assert c["nf2"] == 8 * ref1.sum()
        for r, b in enumerate(rl1):
            assert fof1[b] == 8 * r
            assert np.array_equal(l2pl1[8 * r:8 * r + 8], [b] * 8) and np.array_equal(l2oct[8 * r:8 * r + 8], range(8))
        # 2:1 balance: refined level-1 tiles have all 6 face neighbours present (periodic)
        for b in np.flatnonzero(ref1):
            x, y, z = pos3(f1pos[b], l1gx)
            for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                q = (((z + d[2]) % l1gx) * l1gx + (y + d[1]) % l1gx) * l1gx + (x + d[0]) % l1gx
                assert l1grid[q] >= 0
    # skins: refined or a refined 26-neighbour, ascending
    skin0 = a.table("skin0")
    brute = [cb for cb in range(c["cnb0"]) if any(
        ref0[(((pos3(cb, cgx)[2] + dz) % cgx) * cgx + (pos3(cb, cgx)[1] + dy) % cgx) * cgx + (pos3(cb, cgx)[0] + dx) % cgx]
        for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1))]
    assert list(skin0) == brute


def test_all_all_tables():
    a = orc.Amr(16, l1=orc.MODE_ALL, l2=orc.MODE_ALL)
    assert a.init() == 0
    c = a.counts()
    assert (c["nref0"], c["nf1"], c["nref1"], c["nf2"]) == (8, 64, 64, 512)
    check_table_invariants(a, 16)


def test_slab_all_fixture():
    # Appendix: --l1 slab --l2 all refines coarse z-tiles < cgx/2, and level-2 covers the level-1
    # positions with 1 <= pz <= cgx - 2 (pz = 0 and cgx - 1 are not deep under periodic wrap)
    n0 = 32
    cgx, l1gx = n0 // 8, n0 // 4
    a = orc.Amr(n0, l1=orc.MODE_SLAB, l2=orc.MODE_ALL)
    assert a.init() == 0
    c = a.counts()
    assert c["nref0"] == cgx * cgx * (cgx // 2)
    f1pos, ref1 = a.table("f1pos"), a.table("ref1")
    pz = f1pos // (l1gx * l1gx)
    assert np.array_equal(ref1.astype(bool), (pz >= 1) & (pz <= cgx - 2))
    check_table_invariants(a, n0)


def test_threshold_marks_with_ties():
    n0 = 24
    cnb0 = (n0 // 8) ** 3
    rng = np.random.default_rng(11)
    sen0 = rng.integers(0, 6, cnb0).astype(np.float32)  # heavy ties
    a = orc.Amr(n0, re=1600, l1=orc.MODE_SENSOR, l2=orc.MODE_SENSOR, frac1=0.3, frac2=0.2)
    assert a.init(sen0=sen0) == 0
    tgt = min(a.counts()["cap_nref0"], max(1, int(0.3 * cnb0)))
    thr = np.sort(sen0)[cnb0 - tgt]
    assert np.array_equal(a.table("ref0"), (sen0 >= thr).astype(np.int32))
    assert a.table("ref0").sum() >= tgt
    check_table_invariants(a, n0)


def test_rest_state_is_a_fixed_point_of_the_subcycle():
    n0 = 16
    a = orc.Amr(n0, l1=orc.MODE_SLAB, l2=orc.MODE_ALL, prec=orc.PREC_FP64)
    a.init()
    for lvl in range(3):
        nt = a.raw(lvl).size // (19 * 512)
        a.set_raw(lvl, np.broadcast_to(W[None, :, None], (nt, 19, 512)).ravel())
    a.step(2)
    for lvl in range(3):
        nt = a.raw(lvl).size // (19 * 512)
        assert np.allclose(a.raw(lvl).reshape(nt, 19, 512), W[None, :, None], rtol=0, atol=1e-14)


def test_static_gold_initial_state_is_close_to_tgv():
    # full refinement: level 2 is the 4 N0 lattice seeded by prolongation, within interpolation error
    a = orc.Amr(16, re=800)
    a.init()
    E, ens, mass = a.diag()
    assert E == pytest.approx(0.05 ** 2 / 8, rel=0.15)
    assert mass == pytest.approx(1.0, abs=1e-5)


def test_regrid_with_same_sensor_keeps_tables_and_copies_tiles():
    n0 = 24
    cnb0, ng1 = (n0 // 8) ** 3, (n0 // 4) ** 3
    rng = np.random.default_rng(12)
    sen0 = rng.random(cnb0).astype(np.float32)
    sen1 = rng.random(ng1).astype(np.float32)
    a = orc.Amr(n0, re=1600, l1=orc.MODE_SENSOR, l2=orc.MODE_SENSOR, frac1=0.5, frac2=0.5)
    assert a.init(sen0, sen1) == 0
    a.step(2)
    before = {t: a.table(t) for t in ("ref0", "l1pcb", "f1pos", "ref1", "l2pl1")}
    r1, r2 = a.raw(1), a.raw(2)
    assert a.regrid(sen0, sen1) == 0
    for t, v in before.items():
        assert np.array_equal(a.table(t), v), t
    assert np.array_equal(a.raw(1), r1) and np.array_equal(a.raw(2), r2)


PB = 19 * 4 * (2 * 512 + 1000) * 6 / 5  # pool bytes per fine tile as the budget counts them


def test_overflow_is_reported():
    n0 = 24
    cnb0 = (n0 // 8) ** 3
    cap_nref0 = int(np.ceil(0.1 * cnb0 * 1.4))
    # a budget that holds only the target-sized level-1 pool: all tied marks every tile, beyond it
    a = orc.Amr(n0, re=1600, l1=orc.MODE_SENSOR, l2=orc.MODE_SENSOR, frac1=0.1, frac2=0.1,
                budget_bytes=8 * cap_nref0 * PB * 1.0001)
    assert a.counts()["cap1"] == 8 * cap_nref0
    assert a.init(sen0=np.ones(cnb0, np.float32)) == 3


def test_level1_pool_grows_on_tied_init():
    n0 = 24
    cnb0 = (n0 // 8) ** 3
    a = orc.Amr(n0, re=1600, l1=orc.MODE_SENSOR, l2=orc.MODE_SENSOR, frac1=0.1, frac2=0.1)
    cap1 = a.counts()["cap1"]
    assert a.init(sen0=np.ones(cnb0, np.float32)) == 0  # all tied: every tile marked
    c = a.counts()
    assert c["nf1"] == 8 * cnb0 > cap1 and c["cap1"] == 8 * cnb0 and c["cap_nref0"] == cap1 // 8
    check_table_invariants(a, n0)
    a.step(2)
    assert np.all(np.isfinite(a.diag()))


def test_level1_pool_grows_on_tied_regrid_and_keeps_the_data():
    """Growth during a regrid: tiles that stay refined keep their level-1 populations (the migration
    reads the old pool), new ones are prolonged from level 0."""
    n0 = 32
    cnb0, ng1 = (n0 // 8) ** 3, (n0 // 4) ** 3
    rng = np.random.default_rng(5)
    s0, s1 = rng.random(cnb0).astype(np.float32), rng.random(ng1).astype(np.float32)
    a = orc.Amr(n0, re=1600, l1=orc.MODE_SENSOR, l2=orc.MODE_SENSOR, frac1=0.1, frac2=0.3)
    assert a.init(s0, s1) == 0
    a.step(2)
    cap1 = a.counts()["cap1"]
    old_pos, old_f1 = a.table("f1pos"), a.raw(1)
    assert a.regrid(np.ones(cnb0, np.float32), s1) == 0
    c = a.counts()
    assert c["nf1"] == 8 * cnb0 > cap1 and c["cap1"] >= c["nf1"]
    check_table_invariants(a, n0)
    new_pos, new_f1 = a.table("f1pos"), a.raw(1)
    blk = 19 * 512
    where = {int(p): b for b, p in enumerate(new_pos)}
    for b, p in enumerate(old_pos):
        nb = where[int(p)]
        assert np.array_equal(new_f1[nb * blk:(nb + 1) * blk], old_f1[b * blk:(b + 1) * blk])
    a.step(2)
    assert np.all(np.isfinite(a.diag()))


def test_level2_pool_grows_within_budget_and_overflows_beyond():
    n0 = 32
    cnb0, ng1 = (n0 // 8) ** 3, (n0 // 4) ** 3
    rng = np.random.default_rng(23)
    s0, s1 = rng.random(cnb0).astype(np.float32), rng.random(ng1).astype(np.float32)
    a = orc.Amr(n0, re=1600, l1=orc.MODE_SENSOR, l2=orc.MODE_SENSOR, frac1=0.5, frac2=0.05)
    assert a.init(s0, s1) == 0
    cap2 = a.counts()["cap2"]
    assert a.regrid(s0, np.ones(ng1, np.float32)) == 0
    c = a.counts()
    assert c["nf2"] > cap2 and c["cap2"] >= c["nf2"]
    check_table_invariants(a, n0)
    # with a budget that only fits the initial pools, the same regrid is an overflow
    b = orc.Amr(n0, re=1600, l1=orc.MODE_SENSOR, l2=orc.MODE_SENSOR, frac1=0.5, frac2=0.05,
                budget_bytes=(a.counts()["cap1"] + cap2) * PB * 1.0001)
    assert b.init(s0, s1) == 0
    assert b.regrid(s0, np.ones(ng1, np.float32)) == 3
