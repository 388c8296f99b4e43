"""Independent pins of the oracle's coarse-fine interface (src/tgv3d_amr.cu:56-202, 241-430, 442-453).

Every pin feeds the oracle a state whose correct result follows from mathematics, not from the
oracle's own formulas, and each one fails under a plausible mistake (tests/test_oracle_mutants.py plants
those mistakes in copies of the oracle and checks that the named pins catch them):

* An affine density rho(X) and an affine amplitude eps(X) of a traceless second-order Hermite mode m
  (m_q = w_q Q_q:A, with Q_q = c_q c_q - I/3 and tr A = 0) have zero velocity, so every regularised
  non-equilibrium part is exactly eps m, and trilinear interpolation reproduces affine functions exactly.
  Seeding, ghost fill, migration and restriction must therefore return w rho(X) + s eps(X) m at the target
  cell centre X, with s the level's non-equilibrium scale (nqs, or nqi for restriction: amr.cu:659-663).
  A swapped 1/4-3/4 weight, a ghost or child offset of one cell, or a dropped nqs/nqi breaks that.
* A constant field must be reproduced exactly where a level-2 interpolation drops absent level-1 corners
  (weights renormalised, amr.cu:105, 117, 133-144).
* With omega = 0 a fine step is a pure shift: the pulled value of direction q at cell x is the value at
  x - c_q, taken from the same level where a tile of that level covers x - c_q and from the ghost box
  otherwise (amr.cu:255-266, 340-348). Coded values (distinct for fine data and ghosts) make every read
  position visible, checked by brute force over all (tile, q, cell).
* The sensor of f = feq(rho, u) + eps m is eps |m| / rho (amr.cu:442-453), per cell, maximised per tile.

Geometry: coarse lattice N0 = 32 (4^3 coarse tiles), the 2^3 central coarse tiles refined, level 2 on the
8 deep level-1 tiles. No interpolation stencil crosses the periodic wrap, so the affine closed forms hold.

Run as a script (python tests/interface_pins.py) it prints a JSON object {pin: error or null}; the
oracle build it checks can be redirected with ORACLE_LIB.
"""
import json
import os
import sys

import numpy as np

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1])
C = np.stack([CX, CY, CZ], 1)
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
MB = 8
N0 = 32

# affine density and mode amplitude over coarse-cell coordinates X (coarse cell i has X = i)
RHO_A = np.array([0.0040, 0.0025, 0.0015])
EPS_0, EPS_A = 0.01, np.array([2.0e-4, -1.5e-4, 1.0e-4])
A_MODE = np.array([[0.3, 0.7, -0.2], [0.7, -0.5, 0.4], [-0.2, 0.4, 0.2]])  # symmetric, traceless
QT = np.einsum("qa,qb->qab", C, C) - np.eye(3) / 3
M = W * np.einsum("qab,ab->q", QT, A_MODE)  # traceless Hermite mode: zero mass and momentum
U_A = np.array([0.02, -0.01, 0.015])  # velocity of the sensor pin

CELL = np.stack([np.arange(512) % 8, (np.arange(512) // 8) % 8, np.arange(512) // 64], 1)  # (512, 3) i, j, k
BOX = np.stack([np.arange(1000) % 10, (np.arange(1000) // 10) % 10, np.arange(1000) // 100], 1) - 1
INTERIOR = np.all((BOX >= 0) & (BOX < MB), 1)


def _pos3(p, g):
    return np.stack([p % g, (p // g) % g, p // (g * g)], -1)


def tile_origins(a, level):
    """Origin (in the level's own cell coordinates) of each tile of a level, from the block tables."""
    n0 = a.n
    if level == 0:
        cgx = n0 // MB
        return _pos3(np.arange(cgx ** 3), cgx) * MB
    f1pos = a.table("f1pos")
    p1 = _pos3(f1pos, n0 // 4)
    if level == 1:
        return p1 * MB
    pl1, oct_ = a.table("l2pl1"), a.table("l2oct")
    q = np.stack([oct_ & 1, (oct_ >> 1) & 1, oct_ >> 2], 1)
    return p1[pl1] * 2 * MB + q * MB


def to_coarse(g, level):
    """Coarse-cell coordinate of the centre of a level-`level` cell g."""
    return g / 2.0 ** level + (0.5 / 2.0 ** level - 0.5)


def closed_form(X, s, u=None):
    """(..., 19) populations w rho(X) + s eps(X) m, or feq(rho, u) + s eps m when u is given."""
    rho = 1.0 + X @ RHO_A
    eps = EPS_0 + X @ EPS_A
    if u is None:
        base = W * rho[..., None]
    else:
        cu = 3.0 * (u @ C.T)
        uu = np.sum(u * u, -1)
        base = W * rho[..., None] * (1 + cu + 0.5 * cu * cu - 1.5 * uu[..., None])
    return base + s * eps[..., None] * M


def level_field(a, level, s, u_fn=None):
    """Closed-form state of a whole level in tile-major layout (tiles, 19, 512), and its coordinates."""
    g = tile_origins(a, level)[:, None, :] + CELL[None]
    X = to_coarse(g, level)
    u = None if u_fn is None else u_fn(X)
    f = closed_form(X, s, u)  # (tiles, 512, 19)
    return np.ascontiguousarray(f.transpose(0, 2, 1)), X


def set_level(a, level, f):
    a.set_raw(level, f.ravel())


def get_level(a, level):
    return a.f64(level).reshape(-1, 19, 512)


def tol(a):
    import oracle as orc
    return 1e-13 if a.prec == orc.PREC_FP64 else 3e-6


def build_central(orc, prec=None, extra=None):
    """N0 = 32 grid with the central 2^3 coarse tiles refined and level 2 on the 8 deep level-1 tiles
    (sensors injected; sensor-mode init as the reference runs it, amr.cu:779-815)."""
    prec = orc.PREC_FP64 if prec is None else prec
    cgx = N0 // MB
    t = _pos3(np.arange(cgx ** 3), cgx)
    sel = np.all((t >= 1) & (t <= 2), 1)
    if extra is not None:
        sel &= ~np.all(t == extra, 1)
    sen0 = sel.astype(np.float32)
    # frac1 = refined share exactly, so the order-statistic threshold is the injected value 1
    a = orc.Amr(N0, re=800, l1=orc.MODE_SENSOR, l2=orc.MODE_SENSOR, frac1=sel.sum() / cgx ** 3, frac2=1.0,
                prec=prec)
    assert a.init(sen0, np.ones((N0 // 4) ** 3, np.float32)) == 0
    return a


def params(a):
    p = a.params()
    # the reference's scale factors (amr.cu:659-663), from the level relaxation times; the kernels take
    # them as float arguments (amr.cu:70, 94, 160, 275)
    t0, t1, t2 = p["tau0"], p["tau1"], p["tau2"]
    v = {"nqs01": (t1 - 1) / (2 * (t0 - 1)), "nqi10": 2 * (t0 - 1) / (t1 - 1),
         "nqs12": (t2 - 1) / (2 * (t1 - 1)), "nqi21": 2 * (t1 - 1) / (t2 - 1)}
    return {k: float(np.float32(x)) for k, x in v.items()}


def _close(got, want, atol, what):
    err = np.abs(got - want)
    if not np.all(err <= atol):
        i = np.unravel_index(np.argmax(err), err.shape)
        raise AssertionError(f"{what}: max error {err.max():.3e} at {i} (got {got[i]!r}, want {want[i]!r})")


# ---- pins -----------------------------------------------------------------------------------------

def pin_seed_l1(a):
    """Seeding level 1 from level 0 (prolong_L0, amr.cu:70-91, 299-308)."""
    s = params(a)
    f0, _ = level_field(a, 0, 1.0)
    set_level(a, 0, f0)
    a.run(a.OP_SEED_L1)
    want, _ = level_field(a, 1, s["nqs01"])
    _close(get_level(a, 1), want, tol(a), "seed L1")


def pin_seed_l2(a):
    """Seeding level 2 from level 1 (prolong_L1, amr.cu:94-118, 382-393)."""
    s = params(a)
    f1, _ = level_field(a, 1, 1.0)
    set_level(a, 1, f1)
    a.run(a.OP_SEED_L2)
    want, _ = level_field(a, 2, s["nqs12"])
    _close(get_level(a, 2), want, tol(a), "seed L2")


def _ghost_expect(a, level, s):
    """Expected ghost boxes: the closed form at every box position the fill writes, NaN elsewhere
    (interior positions, and positions whose source block is refined to the fine level: amr.cu:182, 198)."""
    n0 = a.n
    nl = n0 * 2 ** level
    org = tile_origins(a, level)
    g = (org[:, None, :] + BOX[None]) % nl  # (tiles, 1000, 3)
    if level == 1:
        cgx = n0 // MB
        cb = g // (2 * MB)
        skip = a.table("ref0")[(cb[..., 2] * cgx + cb[..., 1]) * cgx + cb[..., 0]] != 0
    else:
        l1gx = n0 // 4
        p = g // (2 * MB)
        lb = a.table("l1grid")[(p[..., 2] * l1gx + p[..., 1]) * l1gx + p[..., 0]]
        ref1 = np.append(a.table("ref1"), 0)
        skip = (lb >= 0) & (ref1[np.where(lb >= 0, lb, -1)] != 0)
    skip |= INTERIOR[None]
    want = closed_form(to_coarse(g, level), s).transpose(0, 2, 1).copy()  # (tiles, 19, 1000)
    want[np.broadcast_to(skip[:, None, :], want.shape)] = np.nan
    return want, skip


def _check_ghost(a, level, want, what):
    got = a.ghost(level)
    nan_w, nan_g = np.isnan(want), np.isnan(got)
    if not np.array_equal(nan_w, nan_g):
        bad = np.argwhere(nan_w != nan_g)[0]
        raise AssertionError(f"{what}: written positions differ from the reference's rule at {tuple(bad)}")
    _close(np.where(nan_g, 0, got), np.where(nan_w, 0, want), tol(a), what)


def pin_fill_l1(a):
    """Level-1 ghost fill: prep (amr.cu:160-171) then trilinear from prep0 (amr.cu:122-132, 173-186)."""
    s = params(a)
    f0, _ = level_field(a, 0, 1.0)
    set_level(a, 0, f0)
    a.set_ghost(1, np.full(a.ghost(1).shape, np.nan))
    a.run(a.OP_FILL_L1)
    want, skip = _ghost_expect(a, 1, s["nqs01"])
    assert (~skip).sum() > 0
    _check_ghost(a, 1, want, "fill L1")


def pin_fill_l2(a):
    """Level-2 ghost fill from prep1 over the scattered level 1 (amr.cu:133-144, 187-202)."""
    s = params(a)
    f1, _ = level_field(a, 1, 1.0)
    set_level(a, 1, f1)
    a.set_ghost(2, np.full(a.ghost(2).shape, np.nan))
    a.run(a.OP_FILL_L2)
    want, skip = _ghost_expect(a, 2, s["nqs12"])
    assert (~skip).sum() > 0
    _check_ghost(a, 2, want, "fill L2")


def _restrict(a, level):
    s = params(a)
    scale = s["nqi10"] if level == 1 else s["nqi21"]
    ff, _ = level_field(a, level, 1.0)
    set_level(a, level, ff)
    # the parent level: a different state, so cells the restriction must not touch stay recognisable
    fc = get_level(a, level - 1)
    fc[:] = (2.0 * W)[None, :, None]
    set_level(a, level - 1, fc)
    a.run(a.OP_RESTRICT_L1 if level == 1 else a.OP_RESTRICT_L2)
    want, _ = level_field(a, level - 1, scale)
    got = get_level(a, level - 1)
    if level == 1:
        covered = a.table("ref0") != 0
    else:
        covered = a.table("ref1") != 0
    want[~covered] = (2.0 * W)[None, :, None]
    assert covered.sum() > 0
    _close(got, want, tol(a), f"restrict L{level} -> L{level - 1}")


def pin_restrict_l1(a):
    """Restriction level 1 -> 0: mean of the 8 children, then feq + nqi10 regularize (amr.cu:275-298)."""
    _restrict(a, 1)


def pin_restrict_l2(a):
    """Restriction level 2 -> 1 (amr.cu:358-381)."""
    _restrict(a, 2)


def pin_sensor(a):
    """Sensor: per tile max over cells of sqrt(sum_q (f - feq)^2) / rho (amr.cu:442-453)."""
    mnorm = np.sqrt(np.sum(M * M))
    u_fn = lambda X: U_A * (1.0 + 0.01 * X[..., :1])  # noqa: E731 (non-zero, non-uniform velocity)
    want = {}
    for level in (0, 1):
        f, X = level_field(a, level, 1.0, u_fn)
        set_level(a, level, f)
        rho = 1.0 + X @ RHO_A
        eps = EPS_0 + X @ EPS_A
        want[level] = np.max(np.abs(eps) * mnorm / rho, 1)
    a.run(a.OP_SENSOR)
    for level, name in ((0, "sen0"), (1, "sen1")):
        got = a.table(name).astype(np.float64)
        # FP64: only the float result rounds; FP32: f - feq cancels ~3 digits of the populations
        rel = 1e-6 if tol(a) < 1e-10 else 2e-4
        _close(got / want[level], np.ones_like(got), rel, f"sensor level {level}")


CODE_Q, FLAG = 2.0 ** 24, 2.0 ** 30


def _code(q, g, nl):
    g = g % nl
    return (q + 1) * CODE_Q + ((g[..., 2] * nl + g[..., 1]) * nl + g[..., 0]) + 1.0


def _ghost_read(a, level):
    """omega = 0: the step is the pure shift x - c_q, reading same-level data where a tile of the level
    covers the source cell and the ghost box elsewhere (checked by brute force over every read)."""
    import oracle as orc
    if a.prec != orc.PREC_FP64:
        return  # codes need exact integers up to 2^31
    nl = a.n * 2 ** level
    org = tile_origins(a, level)
    q = np.arange(19)[None, :, None]
    g = org[:, None, :] + CELL[None]  # (tiles, 512, 3)
    set_level(a, level, _code(q, g[:, None], nl))
    gb = org[:, None, :] + BOX[None]
    a.set_ghost(level, _code(q, gb[:, None], nl) + FLAG)
    a.set_omega(level, 0.0)
    a.run(a.OP_STEP_L1 if level == 1 else a.OP_STEP_L2)
    # brute force: is the source cell covered by a tile of this level?
    ng = nl // MB
    present = np.zeros(ng ** 3, bool)
    tp = org // MB
    present[(tp[:, 2] * ng + tp[:, 1]) * ng + tp[:, 0]] = True
    src = (g[:, None, :, :] - C[None, :, None, :]) % nl  # (tiles, 19, 512, 3)
    st = src // MB
    inside = present[(st[..., 2] * ng + st[..., 1]) * ng + st[..., 0]]
    want = _code(q, src, nl) + np.where(inside, 0.0, FLAG)
    got = get_level(a, level)
    assert (~inside).sum() > 0 and inside.sum() > 0
    if not np.array_equal(got, want):
        bad = tuple(np.argwhere(got != want)[0])
        raise AssertionError(f"ghost read L{level}: (tile, q, cell) {bad} read code {got[bad]!r}, want {want[bad]!r}")


def pin_ghost_read_l1(a):
    """Ghost positions read by the level-1 step (amr.cu:255-266)."""
    _ghost_read(a, 1)


def pin_ghost_read_l2(a):
    """Ghost positions read by the level-2 step (amr.cu:340-348)."""
    _ghost_read(a, 2)


CENTRAL_PINS = [pin_seed_l1, pin_seed_l2, pin_fill_l1, pin_fill_l2, pin_restrict_l1, pin_restrict_l2,
                pin_sensor, pin_ghost_read_l1, pin_ghost_read_l2]


def pin_renormalised_fill(orc, prec=None):
    """Absent level-1 corners are dropped and the weights renormalised (amr.cu:133-144): a constant
    field is reproduced at every level-2 ghost position, including those whose stencil meets an absent
    level-1 tile (2:1 balance checks the 6 faces only, so edge and corner neighbours can be absent)."""
    a = build_central(orc, prec, extra=np.array([1, 1, 1]))
    n0, l1gx = a.n, a.n // 4
    const = (W * 1.05 + 0.003 * M)[None, :, None]
    f1 = np.broadcast_to(const, get_level(a, 1).shape)
    set_level(a, 1, np.ascontiguousarray(f1))
    a.set_ghost(2, np.full(a.ghost(2).shape, np.nan))
    a.run(a.OP_FILL_L2)
    s = params(a)["nqs12"]
    got = a.ghost(2)
    # positions whose trilinear stencil meets an absent level-1 cell
    org = tile_origins(a, 2)
    g = (org[:, None, :] + BOX[None]) % (4 * n0)
    base = np.floor((g + 0.5) / 2 - 0.5).astype(int)
    grid = a.table("l1grid")
    absent_any = np.zeros(g.shape[:2], bool)
    for d in np.ndindex(2, 2, 2):
        c = (base + np.array(d)) % (2 * n0)
        p = c // MB
        absent_any |= grid[(p[..., 2] * l1gx + p[..., 1]) * l1gx + p[..., 0]] < 0
    written = ~np.isnan(got[:, 0, :])
    assert np.any(written & absent_any), "geometry has no renormalised ghost position"
    want = (W * 1.05 + 0.003 * s * M)[None, :, None]
    _close(np.where(written[:, None, :], got, want), np.broadcast_to(want, got.shape), tol(a), "renormalised fill L2")


def pin_migrate(orc, prec=None):
    """Regrid migration (amr.cu:397-430): tiles whose parent stays refined are bit copies, new tiles are
    prolonged from level 0 (level 1) or from the migrated level 1 (level 2)."""
    a = build_central(orc, prec)
    s = params(a)
    f0, _ = level_field(a, 0, 1.0)
    set_level(a, 0, f0)
    set_level(a, 1, level_field(a, 1, s["nqs01"])[0])
    # level 2 carries a marker per tile, so a copy from the wrong old tile cannot pass as a copy
    f2 = level_field(a, 2, s["nqs12"] * s["nqs01"])[0]
    f2 += 1e-3 * (np.arange(f2.shape[0]) + 1.0)[:, None, None] * W[None, :, None]
    set_level(a, 2, f2)
    old = {lv: (tile_origins(a, lv), get_level(a, lv)) for lv in (1, 2)}
    cgx = N0 // MB
    t = _pos3(np.arange(cgx ** 3), cgx)
    sen0 = (np.all((t[:, 1:] >= 1) & (t[:, 1:] <= 2), 1) & (t[:, 0] >= 1)).astype(np.float32)
    assert a.regrid(sen0, np.ones((N0 // 4) ** 3, np.float32)) == 0
    assert a.counts()["nf1"] == 8 * 12 and a.counts()["nf2"] == 8 * 16
    for lv, sc in ((1, s["nqs01"]), (2, s["nqs12"] * s["nqs01"])):
        org = tile_origins(a, lv)
        got = get_level(a, lv)
        want, X = level_field(a, lv, sc)
        before = {tuple(o): i for i, o in enumerate(old[lv][0])}
        n_new = 0
        for i, o in enumerate(org):
            j = before.get(tuple(o))
            if j is not None:
                if not np.array_equal(got[i], old[lv][1][j]):
                    raise AssertionError(f"migrate L{lv}: persisting tile at {tuple(o)} is not a bit copy")
                continue
            n_new += 1
            ok = np.all((X[i] >= 0) & (X[i] <= N0 - 1), 1)  # cells whose stencil does not wrap
            _close(got[i][:, ok], want[i][:, ok], tol(a), f"migrate L{lv} new tile at {tuple(o)}")
        assert n_new > 0


ALL_PINS = [p.__name__ for p in CENTRAL_PINS] + ["pin_renormalised_fill", "pin_migrate"]


def run_all(orc, prec=None):
    """{pin name: None if it holds, else the error text}."""
    out = {}
    a = build_central(orc, prec)
    for p in CENTRAL_PINS:
        try:
            p(a)
            out[p.__name__] = None
        except Exception as e:  # noqa: BLE001 (a planted mistake may fail in any way)
            out[p.__name__] = f"{type(e).__name__}: {e}"
    for p in (pin_renormalised_fill, pin_migrate):
        try:
            p(orc, prec)
            out[p.__name__] = None
        except Exception as e:  # noqa: BLE001
            out[p.__name__] = f"{type(e).__name__}: {e}"
    return out


if __name__ == "__main__":
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import oracle

    print(json.dumps(run_all(oracle)))
