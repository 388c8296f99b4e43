"""Physics gates of the reference (src/tgv3d_uniform.cu:197-202; src/tgv3d_amr.cu:899-906;
docs/validation.md:12-14) on the GPU path: the uniform TGV breaks down with a finite enstrophy peak,
monotonic energy and held mass; the static gold (every level refined) reproduces the uniform 4 N0 run
(enstrophy-peak time within 0.3 t*, E/E0 within 2 %)."""
import math

import numpy as np
import pytest

import jpabi

pytestmark = pytest.mark.gpu


def curve_uniform(n, re, tstar, every):
    u = jpabi.Uniform(n, re=re)
    u.init()
    kf = u.derived()["kf"]
    steps = int(tstar / (0.05 * kf) + 0.5)
    out, s = [], 0
    while True:
        out.append((s * 0.05 * kf, *u.diag()))
        if s >= steps:
            break
        k = min(every, steps - s)
        u.step(k)
        s += k
    return np.array(out)


def curve_amr(n0, re, tstar, every):
    a = jpabi.Amr(n0, re=re, l1=jpabi.MODE_ALL, l2=jpabi.MODE_ALL)
    a.init()
    a.use_graph(True)
    kf = a.derived()["kf"]
    steps = int(tstar / (0.05 * kf) + 0.5)
    out, s = [], 0
    while True:
        out.append((s * 0.05 * kf, *a.diag()))
        if s >= steps:
            break
        k = min(every, steps - s)
        a.step(k)
        s += k
    return np.array(out)


def gate(c):
    t, E, ens, m = c.T
    ip = int(np.argmax(ens))
    mono = bool(np.all(E[1:] <= E[:-1] * 1.0000001))
    drift = float(np.max(np.abs(m - m[0]) / m[0]))
    return t[ip], mono, drift


def test_uniform_tgv_gate():
    # docs/validation.md:12: N = 192, Re 1600 peaks in [7.5, 10] (reference 8.93), energy monotonic at
    # the driver's sampling cadence max(1, steps/240)
    n = 192
    steps = int(12.0 / (0.05 * 2 * math.pi / n) + 0.5)
    c = curve_uniform(n, 1600.0, 12.0, max(1, steps // 240))
    tp, mono, drift = gate(c)
    assert math.isfinite(c[:, 2].max())
    assert 7.5 <= tp <= 10.0, tp
    # mass drift: the coded gate (src/tgv3d_uniform.cu:199) is 5e-4. The drift is the FP32 weight-sum
    # bias: the stored weights sum to 1 + 1.49e-8, so each collision adds omega x 1.49e-8 of mass
    # (omega 1.989 here), 2.17e-4 over the 7334 steps to t* = 12 whatever the rounding order; the
    # docs' "< 1e-4" holds only up to about t* = 5.5 of that bias
    assert mono and drift < 5e-4


def test_static_gold_reproduces_uniform_4n():
    # docs/validation.md:13: refine every level at N0 = 64, Re 800 vs uniform 256^3
    n0, re = 64, 800.0
    a = curve_amr(n0, re, 12.0, 16)  # coarse steps; t* per coarse step = 0.05 * 2 pi / n0
    u = curve_uniform(4 * n0, re, 12.0, 64)
    ta, mono_a, drift_a = gate(a)
    tu, mono_u, _ = gate(u)
    assert abs(ta - tu) <= 0.3, (ta, tu)
    assert drift_a < 5e-4
    # E/E0 curves within 2 % at common times
    ea = np.interp(u[:, 0], a[:, 0], a[:, 1] / a[0, 1])
    eu = u[:, 1] / u[0, 1]
    assert np.max(np.abs(ea - eu) / eu) < 0.02
