"""GPU checks of the batched ensemble: every member advanced by the batched launches is bit-identical to
the same member advanced alone (single-context launches, which are themselves bit-exact with the
oracle), with and without hipGraph replay, across regrids; autotuning leaves the state unchanged."""
import numpy as np
import pytest

import jpabi

pytestmark = pytest.mark.gpu

MEMBERS = [
    {"n": 16, "re": 1600.0, "u0": 0.05, "amr": 0},
    {"n": 24, "re": 900.0, "u0": 0.04, "amr": 0, "phase": (0.3, 1.1, 2.0)},
    {"n": 16, "re": 800.0, "u0": 0.05, "amr": 1, "l1": jpabi.MODE_SLAB, "l2": jpabi.MODE_ALL},
    {"n": 24, "re": 1600.0, "u0": 0.06, "amr": 1, "l1": jpabi.MODE_SENSOR, "l2": jpabi.MODE_SENSOR, "frac1": 0.4,
     "frac2": 0.3, "adapt": 3, "phase": (1.0, 0.5, 0.25)},
]


def standalone(m, prec):
    if m["amr"]:
        return jpabi.Amr(m["n"], re=m["re"], u0=m["u0"], l1=m["l1"], l2=m["l2"], frac1=m.get("frac1", 0.3),
                         frac2=m.get("frac2", 0.1), adapt=m.get("adapt", 10), prec=prec, maxgb=16.0,
                         phase=m.get("phase", (0.0, 0.0, 0.0)))
    return jpabi.Uniform(m["n"], re=m["re"], u0=m["u0"], prec=prec, phase=m.get("phase", (0.0, 0.0, 0.0)))


def levels(m):
    return 3 if m["amr"] else 1


def same(ens, alone):
    for m, e, a in zip(MEMBERS, ens.members, alone):
        for lvl in range(levels(m)):
            assert np.array_equal(e.raw(lvl), a.raw(lvl)), (m, lvl)


@pytest.mark.parametrize("prec", [jpabi.PREC_FP32, jpabi.PREC_FP16])
def test_batched_members_match_single_contexts(prec):
    ens = jpabi.Ensemble(MEMBERS, prec=prec, maxgb=16.0)
    ens.init()
    alone = [standalone(m, prec) for m in MEMBERS]
    for a in alone:
        a.init()
    same(ens, alone)  # identical device code paths from the same parameters
    ens.step(3)
    for a in alone:
        a.step(3)
    same(ens, alone)


def test_graph_replay_and_regrid_match():
    prec = jpabi.PREC_FP32
    ens = jpabi.Ensemble(MEMBERS, prec=prec, maxgb=16.0)
    ens.init()
    alone = [standalone(m, prec) for m in MEMBERS]
    for a in alone:
        a.init()
    ens.use_graph(True)
    for _ in range(3):
        ens.step(4)
        for a in alone:
            a.step(4)
        same(ens, alone)
        ens.regrid()
        for m, a in zip(MEMBERS, alone):
            if m["amr"] and (m["l1"] == jpabi.MODE_SENSOR or m["l2"] == jpabi.MODE_SENSOR):
                a.regrid()
        same(ens, alone)


# members whose lattices split into whole 2048-cell chunks (edges that are multiples of 16): the dense step
# can then run 8 cells per lane
# 8 cells per lane takes uniform members whose lattices are multiples of 2048 cells
MEMBERS16 = [
    {"n": 32, "re": 1600.0, "u0": 0.05, "amr": 0},
    {"n": 16, "re": 900.0, "u0": 0.04, "amr": 0, "phase": (0.3, 1.1, 2.0)},
    {"n": 48, "re": 1600.0, "u0": 0.06, "amr": 0, "phase": (1.0, 0.5, 0.25)},
]


@pytest.mark.parametrize("prec", [jpabi.PREC_FP32, jpabi.PREC_FP16])
def test_dense_tile_shapes_match_single_contexts(prec):
    """The batched dense step with 8 cells per lane and with 1 computes what the members compute alone."""
    ens = jpabi.Ensemble(MEMBERS16, prec=prec, maxgb=16.0)
    ens.init()
    alone = [standalone(m, prec) for m in MEMBERS16]
    for a in alone:
        a.init()
    assert ens.dense_shape(8) == 8
    for shape in (8, 1, 8):
        assert ens.dense_shape(shape) == shape
        ens.step(3)
        for a in alone:
            a.step(3)
        for m, e, a in zip(MEMBERS16, ens.members, alone):
            for lvl in range(levels(m)):
                assert np.array_equal(e.raw(lvl), a.raw(lvl)), (shape, m, lvl)


@pytest.mark.parametrize("members", [MEMBERS, MEMBERS16[:2] + [MEMBERS[2]]], ids=["ragged", "amr"])
def test_dense8_needs_uniform_whole_chunks(members):
    """A member whose lattice is not a multiple of 2048 cells (24^3), or an AMR member (whose level 0 is
    tile-major), keeps the dense step at 1 cell per lane."""
    ens = jpabi.Ensemble(members, maxgb=16.0)
    ens.init()
    with pytest.raises(jpabi.JPError) as e:
        ens.dense_shape(8)
    assert e.value.code == jpabi.JP_ERR_ARG
    assert ens.dense_shape(0) == 1


def test_autotune_keeps_state():
    ens = jpabi.Ensemble(MEMBERS, maxgb=16.0)
    ens.init()
    ens.step(2)
    before = [[e.raw(l) for l in range(levels(m))] for m, e in zip(MEMBERS, ens.members)]
    groups = ens.autotune()
    assert len(groups) == jpabi.ENS_KERNELS and all(g > 0 for g in groups)
    after = [[e.raw(l) for l in range(levels(m))] for m, e in zip(MEMBERS, ens.members)]
    for b, a in zip(before, after):
        for x, y in zip(b, a):
            assert np.array_equal(x, y)


def test_ensemble_bench_counts_member_work():
    ens = jpabi.Ensemble(MEMBERS, maxgb=16.0)
    ens.init()
    r = ens.bench(4, warmup=0, graph=True)
    work = 0
    for m, e in zip(MEMBERS, ens.members):
        if m["amr"]:
            c = e.counts()
            work += (c["cnb0"] + 2 * c["nf1"] + 4 * c["nf2"]) * 512
        else:
            work += m["n"] ** 3
    assert r["work"] == pytest.approx(4 * work)
