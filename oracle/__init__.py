"""TEST INFRASTRUCTURE ONLY: ctypes bindings of the CPU oracle (oracle/liboracle.so).

The oracle restates the reference method of src/tgv3d_uniform.cu and src/tgv3d_amr.cu on the CPU. It is
imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker.
Buffers use the reference's tile-major layout f[(b*19+q)*512 + cell].
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB names another build of the same sources: the mutation tests load oracle copies with one
# planted mistake through it, in a child process.
LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(_HERE, "build", "liboracle.so")
_lib = None

PREC_FP32, PREC_FP16, PREC_BF16, PREC_FP64 = 0, 1, 2, 3
MODE_ALL, MODE_SLAB, MODE_SENSOR = 0, 1, 2
RAW_DTYPE = {PREC_FP32: np.float32, PREC_FP16: np.uint16, PREC_BF16: np.uint16, PREC_FP64: np.float64}
COUNT_NAMES = ["cnb0", "nref0", "nf1", "nref1", "nf2", "n_skin0", "n_skin1", "cap_nref0", "cap1",
               "cap_nref1", "cap2", "hw_nf1", "hw_nf2", "overflow"]
TABLES = ["ref0", "fof0", "rl0", "l1pcb", "l1oct", "f1pos", "l1grid", "ref1", "fof1", "rl1", "l2pl1",
          "l2oct", "skin0", "skin1", "sen0", "sen1"]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, i, l, d, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_float
        sig = {
            "orc_threads": (i, []), "orc_set_threads": (None, [i]),
            "orc_f32_to_f16": (None, [vp, vp, l]), "orc_f16_to_f32": (None, [vp, vp, l]),
            "orc_f32_to_bf16": (None, [vp, vp, l]),
            "orc_feq": (None, [i, f, f, f, f, vp]), "orc_regularize": (None, [vp]),
            "orc_regularize64": (None, [vp]),
            "orc_uni_create": (vp, [i, i, d, d, f]), "orc_uni_destroy": (None, [vp]),
            "orc_uni_init": (None, [vp, vp]), "orc_uni_step": (None, [vp, l]),
            "orc_uni_diag": (None, [vp, vp]), "orc_uni_elem": (l, [vp]),
            "orc_uni_get_raw": (None, [vp, vp]), "orc_uni_set_raw": (None, [vp, vp]),
            "orc_uni_get_f64": (None, [vp, vp]), "orc_uni_set_omega": (None, [vp, d]), "orc_uni_time_steps": (d, [vp, l]),
            "orc_amr_create": (vp, [vp, vp]), "orc_amr_destroy": (None, [vp]),
            "orc_amr_init": (i, [vp, vp, vp]), "orc_amr_regrid": (i, [vp, vp, vp]),
            "orc_amr_step": (None, [vp, l]), "orc_amr_diag": (None, [vp, vp]),
            "orc_amr_counts": (None, [vp, vp]), "orc_amr_table": (i, [vp, i, vp, l]),
            "orc_amr_level_elems": (l, [vp, i]), "orc_amr_elem": (l, [vp]),
            "orc_amr_get_raw": (None, [vp, i, vp]), "orc_amr_set_raw": (None, [vp, i, vp]),
            "orc_amr_get_f64": (None, [vp, i, vp]), "orc_amr_params": (None, [vp, vp]),
            "orc_uni_qspeed": (None, [vp, vp, vp, vp]), "orc_amr_qspeed": (None, [vp, vp, vp]),
            "orc_qspeed_field": (None, [vp, i, i, vp, vp, vp]),
            "orc_amr_run": (None, [vp, i]), "orc_amr_set_omega": (None, [vp, i, d]),
            "orc_amr_ghost_elems": (l, [vp, i]), "orc_amr_get_ghost": (None, [vp, i, vp]),
            "orc_amr_set_ghost": (None, [vp, i, vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def qspeed_field(F, diff_in_field=True, velocity=True):
    """Q, speed (and velocity) of an (N, N, N, 4) float32 field of (rho, u, v, w), index [k, j, i]."""
    F = np.ascontiguousarray(F, dtype=np.float32)
    n = F.shape[0]
    q = np.empty(n ** 3, np.float32)
    s = np.empty(n ** 3, np.float32)
    v = np.empty(3 * n ** 3, np.float32) if velocity else None
    lib().orc_qspeed_field(_ptr(F), n, 1 if diff_in_field else 0, _ptr(q), _ptr(s), _ptr(v) if velocity else None)
    out = {"Q": q.reshape(n, n, n), "speed": s.reshape(n, n, n)}
    if velocity:
        out["velocity"] = v.reshape(n, n, n, 3)
    return out


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def default_scale(u0):
    """Power-of-two scale of the shifted FP16/BF16 store: the largest 2^k with 2^k * u0 <= 2^11."""
    import math
    return float(2.0 ** math.floor(math.log2(2048.0 / u0)))


class Uniform:
    """Uniform periodic TGV lattice (src/tgv3d_uniform.cu:142-195) on the CPU."""

    def __init__(self, n, re=1600.0, u0=0.05, prec=PREC_FP32, scale=None):
        self.n, self.re, self.u0, self.prec = n, re, u0, prec
        self.scale = float(scale if scale is not None else (default_scale(u0) if prec in (1, 2) else 1.0))
        self.h = lib().orc_uni_create(prec, n, re, u0, self.scale)
        if not self.h:
            raise ValueError("bad precision")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_uni_destroy(self.h)
            self.h = None

    def init(self, phase=None):
        ph = None if phase is None else np.ascontiguousarray(phase, dtype=np.float64)
        lib().orc_uni_init(self.h, _ptr(ph))

    def step(self, n=1):
        lib().orc_uni_step(self.h, n)

    def field(self):
        """Q, speed, velocity of the output field (the uniform writer's arrays), index [k, j, i]."""
        n = self.n
        q, s, v = np.empty(n ** 3, np.float32), np.empty(n ** 3, np.float32), np.empty(3 * n ** 3, np.float32)
        lib().orc_uni_qspeed(self.h, _ptr(q), _ptr(s), _ptr(v))
        return {"Q": q.reshape(n, n, n), "speed": s.reshape(n, n, n), "velocity": v.reshape(n, n, n, 3)}

    def diag(self):
        out = np.zeros(3)
        lib().orc_uni_diag(self.h, _ptr(out))
        return out

    def raw(self):
        out = np.empty(self.n ** 3 * 19, dtype=RAW_DTYPE[self.prec])
        lib().orc_uni_get_raw(self.h, _ptr(out))
        return out

    def set_raw(self, a):
        a = np.ascontiguousarray(a, dtype=RAW_DTYPE[self.prec])
        assert a.size == self.n ** 3 * 19
        lib().orc_uni_set_raw(self.h, _ptr(a))

    def f64(self):
        out = np.empty(self.n ** 3 * 19, dtype=np.float64)
        lib().orc_uni_get_f64(self.h, _ptr(out))
        return out

    def set_omega(self, w):
        lib().orc_uni_set_omega(self.h, w)

    def time_steps(self, n):
        return lib().orc_uni_time_steps(self.h, n)


class Amr:
    """3-level block AMR (src/tgv3d_amr.cu:643-837) on the CPU."""

    def __init__(self, n=32, re=800.0, u0=0.05, l1=MODE_ALL, l2=MODE_ALL, frac1=0.35, frac2=0.35,
                 prec=PREC_FP32, budget_bytes=1e30, skipprep=0, skiprestr=0, phase=(0, 0, 0), scales=None,
                 develop=300):
        """develop: hidden level-0 steps of a sensor-mode initial build (the reference's 300); a test that
        copies a developed state in from elsewhere can skip them."""
        self.n, self.prec = n, prec
        if scales is None:
            s = default_scale(u0) if prec in (1, 2) else 1.0
            scales = (s, s, s)
        self.scales = tuple(float(x) for x in scales)
        ip = np.array([n, l1, l2, skipprep, skiprestr, prec, develop], dtype=np.int32)
        dp = np.array([re, u0, frac1, frac2, budget_bytes, *phase, *self.scales], dtype=np.float64)
        self.h = lib().orc_amr_create(_ptr(ip), _ptr(dp))
        if not self.h:
            raise ValueError("bad precision")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_amr_destroy(self.h)
            self.h = None

    def init(self, sen0=None, sen1_by_pos=None):
        s0 = None if sen0 is None else np.ascontiguousarray(sen0, dtype=np.float32)
        s1 = None if sen1_by_pos is None else np.ascontiguousarray(sen1_by_pos, dtype=np.float32)
        return lib().orc_amr_init(self.h, _ptr(s0), _ptr(s1))

    def regrid(self, sen0=None, sen1_by_pos=None):
        s0 = None if sen0 is None else np.ascontiguousarray(sen0, dtype=np.float32)
        s1 = None if sen1_by_pos is None else np.ascontiguousarray(sen1_by_pos, dtype=np.float32)
        return lib().orc_amr_regrid(self.h, _ptr(s0), _ptr(s1))

    def step(self, n=1):
        lib().orc_amr_step(self.h, n)

    def field(self):
        """Q and speed of the composite at 4 N0 (the AMR writer's arrays), index [k, j, i]."""
        n = 4 * self.n
        q, s = np.empty(n ** 3, np.float32), np.empty(n ** 3, np.float32)
        lib().orc_amr_qspeed(self.h, _ptr(q), _ptr(s))
        return {"Q": q.reshape(n, n, n), "speed": s.reshape(n, n, n)}

    def diag(self):
        out = np.zeros(3)
        lib().orc_amr_diag(self.h, _ptr(out))
        return out

    def counts(self):
        out = np.zeros(14, dtype=np.int64)
        lib().orc_amr_counts(self.h, _ptr(out))
        return dict(zip(COUNT_NAMES, (int(x) for x in out)))

    def table(self, name):
        which = TABLES.index(name)
        cap = 1 << 24
        dt = np.float32 if name.startswith("sen") else np.int32
        buf = np.empty(cap, dtype=dt)
        n = lib().orc_amr_table(self.h, which, _ptr(buf), cap)
        if n < 0:
            raise ValueError(name)
        return buf[:n].copy()

    def raw(self, level):
        n = lib().orc_amr_level_elems(self.h, level)
        out = np.empty(n, dtype=RAW_DTYPE[self.prec])
        lib().orc_amr_get_raw(self.h, level, _ptr(out))
        return out

    def set_raw(self, level, a):
        n = lib().orc_amr_level_elems(self.h, level)
        a = np.ascontiguousarray(a, dtype=RAW_DTYPE[self.prec])
        assert a.size == n
        lib().orc_amr_set_raw(self.h, level, _ptr(a))

    def f64(self, level):
        n = lib().orc_amr_level_elems(self.h, level)
        out = np.empty(n, dtype=np.float64)
        lib().orc_amr_get_f64(self.h, level, _ptr(out))
        return out

    # single operations of the subcycle / regrid (oracle/capi.cpp AmrO::run)
    OP_FILL_L1, OP_FILL_L2, OP_STEP_L1, OP_STEP_L2, OP_RESTRICT_L1, OP_RESTRICT_L2, OP_SEED_L1, OP_SEED_L2, \
        OP_SENSOR = range(9)

    def run(self, op):
        lib().orc_amr_run(self.h, op)

    def set_omega(self, level, w):
        lib().orc_amr_set_omega(self.h, level, w)

    def ghost(self, level):
        """Ghost boxes of a fine level, (tiles, 19, 1000), box index (ek*10 + ej)*10 + ei."""
        n = lib().orc_amr_ghost_elems(self.h, level)
        out = np.empty(n, dtype=np.float64)
        lib().orc_amr_get_ghost(self.h, level, _ptr(out))
        return out.reshape(-1, 19, 1000)

    def set_ghost(self, level, a):
        n = lib().orc_amr_ghost_elems(self.h, level)
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        assert a.size == n
        lib().orc_amr_set_ghost(self.h, level, _ptr(a))

    def params(self):
        out = np.zeros(12)
        lib().orc_amr_params(self.h, _ptr(out))
        names = ["nu", "tau0", "tau1", "tau2", "om0", "om1", "om2", "nqs01", "nqi10", "nqs12", "nqi21", "kf"]
        return dict(zip(names, out))
