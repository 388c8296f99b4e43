"""ctypes binding of the C ABI in include/jp.h (build/libjp.so), for the tests, bench.py and scripts.

The library is the product: there is no Python or CPU fallback. If build/libjp.so is missing, loading
fails with an error that says how to build it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libjp.so")

JP_OK, JP_ERR_HIP, JP_ERR_ARG, JP_ERR_CAPACITY, JP_ERR_STATE, JP_ERR_NOMEM = range(6)
PREC_FP32, PREC_FP16, PREC_BF16, PREC_FP64 = 0, 1, 2, 3
PREC_NAMES = {"fp32": 0, "fp16": 1, "bf16": 2, "fp64": 3}
MODE_ALL, MODE_SLAB, MODE_SENSOR = 0, 1, 2
MODE_NAMES = {"all": 0, "slab": 1, "sensor": 2}
RAW_DTYPE = {PREC_FP32: np.float32, PREC_FP16: np.uint16, PREC_BF16: np.uint16, PREC_FP64: np.float64}

_lib = None


class JPError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"jp error {code}: {msg}")
        self.code = code


class UniformParams(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("re", ctypes.c_double), ("u0", ctypes.c_double), ("prec", ctypes.c_int),
                ("phase", ctypes.c_double * 3), ("scale", ctypes.c_double), ("device", ctypes.c_int),
                ("plane_pad", ctypes.c_long)]


class AmrParams(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("re", ctypes.c_double), ("u0", ctypes.c_double), ("prec", ctypes.c_int),
                ("phase", ctypes.c_double * 3), ("l1_mode", ctypes.c_int), ("l2_mode", ctypes.c_int),
                ("frac1", ctypes.c_double), ("frac2", ctypes.c_double), ("adapt", ctypes.c_int),
                ("maxgb", ctypes.c_double), ("skipprep", ctypes.c_int), ("skiprestr", ctypes.c_int),
                ("scale", ctypes.c_double * 3), ("device", ctypes.c_int)]


ENS_KERNELS = 7  # JP_ENS_KERNELS
COUNT_NAMES = ["cnb0", "nref0", "nf1", "nref1", "nf2", "n_skin0", "n_skin1", "cap_nref0", "cap1", "cap_nref1",
               "cap2", "hw_nf1", "hw_nf2", "ndeep"]


class Counts(ctypes.Structure):
    _fields_ = [(k, ctypes.c_long) for k in COUNT_NAMES]

    def as_dict(self):
        return {k: getattr(self, k) for k in COUNT_NAMES}


TABLES = ["ref0", "fof0", "rl0", "l1pcb", "l1oct", "f1pos", "l1grid", "ref1", "fof1", "rl1", "l2pl1", "l2oct",
          "skin0", "skin1", "sen0", "sen1"]


class Derived(ctypes.Structure):
    _fields_ = [("nu", ctypes.c_double), ("tau0", ctypes.c_double), ("tau1", ctypes.c_double),
                ("tau2", ctypes.c_double), ("omega0", ctypes.c_double), ("omega1", ctypes.c_double),
                ("omega2", ctypes.c_double), ("kf", ctypes.c_double), ("nqs01", ctypes.c_double),
                ("nqi10", ctypes.c_double), ("nqs12", ctypes.c_double), ("nqi21", ctypes.c_double),
                ("scale", ctypes.c_double * 3), ("prec", ctypes.c_int), ("n", ctypes.c_int),
                ("cells", ctypes.c_long)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["scale"] = list(self.scale)
        return d


class BenchResult(ctypes.Structure):
    _fields_ = [("ms_total", ctypes.c_double), ("ms_per_step", ctypes.c_double), ("work", ctypes.c_double),
                ("glups", ctypes.c_double), ("bytes_per_update", ctypes.c_double), ("gbps", ctypes.c_double),
                ("pool_gb", ctypes.c_double), ("steps", ctypes.c_long), ("device_used_gb", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):  # nothing built yet: build the gfx950 library (hipcc) once
            try:
                r = subprocess.run(["make", "-j", str(min(8, os.cpu_count() or 1)), "-C", HERE], capture_output=True,
                                   text=True)
                log = r.stdout[-2000:] + r.stderr[-2000:]
            except OSError as e:
                log = str(e)
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(f"{LIB_PATH} is missing and `make -C {HERE}` did not build it; there is no "
                                   f"fallback path:\n{log}")
        L = ctypes.CDLL(LIB_PATH)
        vp, i, l, d, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_size_t
        sig = {
            "jp_version": (ctypes.c_char_p, []),
            "jp_uniform_create": (i, [ctypes.POINTER(UniformParams), ctypes.POINTER(vp)]),
            "jp_init": (i, [vp]), "jp_step": (i, [vp, l]), "jp_diag": (i, [vp, vp]),
            "jp_diag_noalloc": (i, [vp, vp]),
            "jp_bench": (i, [vp, l, l, i, ctypes.POINTER(BenchResult)]),
            "jp_get_derived": (i, [vp, ctypes.POINTER(Derived)]),
            "jp_get_populations": (i, [vp, i, vp, sz]), "jp_set_populations": (i, [vp, i, vp, sz]),
            "jp_population_bytes": (i, [vp, i, ctypes.POINTER(sz)]),
            "jp_sync": (i, [vp]), "jp_last_error": (ctypes.c_char_p, [vp]), "jp_destroy": (None, [vp]),
            "jp_probe_copy": (i, [i, sz, i, ctypes.POINTER(d), ctypes.POINTER(d)]),
            "jp_probe_bandwidth": (i, [i, i, sz, i, ctypes.POINTER(d)]),
            "jp_field_edge": (ctypes.c_long, [vp]), "jp_field": (i, [vp, vp, vp, vp, ctypes.c_long]),
            "jp_write_vti": (i, [vp, ctypes.c_char_p]), "jp_write_boxes": (i, [vp, ctypes.c_char_p]),
            "jp_amr_create": (i, [ctypes.POINTER(AmrParams), ctypes.POINTER(vp)]),
            "jp_regrid": (i, [vp]), "jp_init_with_sensor": (i, [vp, vp, vp]),
            "jp_regrid_with_sensor": (i, [vp, vp, vp]),
            "jp_get_counts": (i, [vp, ctypes.POINTER(Counts)]),
            "jp_get_table": (i, [vp, i, vp, l, ctypes.POINTER(l)]),
            "jp_use_graph": (i, [vp, i]), "jp_set_variant": (i, [vp, i]),
            "jp_set_step_pairs": (i, [vp, i]), "jp_step_pair_shape": (i, [vp, vp]),
            "jp_set_step_pair_shape": (i, [vp, i]), "jp_steps_per_launch": (i, [vp, ctypes.POINTER(i)]),
            "jp_ensemble_dense_shape": (i, [vp, i, ctypes.POINTER(i)]), "jp_ensemble_create": (i, [vp, i, i, d, i, ctypes.POINTER(vp)]),
            "jp_ensemble_member": (i, [vp, i, ctypes.POINTER(vp)]), "jp_autotune": (i, [vp, vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def declared_symbols(header=None):
    """Function names declared in include/jp.h."""
    import re
    header = header or os.path.join(HERE, "..", "include", "jp.h")
    txt = open(header).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+\**(jp_\w+)\s*\(", txt, re.M)))


def probe_bandwidth(nbytes, iters=10, device=0):
    """Plain-kernel HBM bandwidth in GB/s: {"copy": read + write, "read": ..., "write": ...}."""
    out = {}
    for mode, name in enumerate(("copy", "read", "write")):
        g = ctypes.c_double()
        rc = lib().jp_probe_bandwidth(device, mode, nbytes, iters, ctypes.byref(g))
        if rc != JP_OK:
            raise JPError(rc, "probe_bandwidth failed")
        out[name] = g.value
    return out


def probe_copy(nbytes, iters=20, device=0):
    g, ms = ctypes.c_double(), ctypes.c_double()
    rc = lib().jp_probe_copy(device, nbytes, iters, ctypes.byref(g), ctypes.byref(ms))
    if rc:
        raise JPError(rc, "probe_copy failed")
    return g.value, ms.value


class _Ctx:
    h = None

    def _check(self, rc):
        if rc != JP_OK:
            msg = lib().jp_last_error(self.h).decode() if self.h else ""
            raise JPError(rc, msg)

    def close(self):
        if self.h:
            lib().jp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def init(self):
        self._check(lib().jp_init(self.h))

    def step(self, n=1):
        self._check(lib().jp_step(self.h, n))

    def diag(self):
        out = np.zeros(3)
        self._check(lib().jp_diag(self.h, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def diag_noalloc(self):
        """jp_diag of a uniform lattice through the allocation-free entry point."""
        out = np.zeros(3)
        self._check(lib().jp_diag_noalloc(self.h, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def field(self, velocity=False):
        """Q-criterion, speed (and velocity) of the output field, (edge, edge, edge) arrays [k, j, i]."""
        n = lib().jp_field_edge(self.h)
        q = np.empty(n ** 3, np.float32)
        s = np.empty(n ** 3, np.float32)
        v = np.empty(3 * n ** 3, np.float32) if velocity else None
        vptr = v.ctypes.data_as(ctypes.c_void_p) if velocity else None
        self._check(lib().jp_field(self.h, q.ctypes.data_as(ctypes.c_void_p), s.ctypes.data_as(ctypes.c_void_p), vptr,
                                   n ** 3))
        out = {"Q": q.reshape(n, n, n), "speed": s.reshape(n, n, n)}
        if velocity:
            out["velocity"] = v.reshape(n, n, n, 3)
        return out

    def write_vti(self, path):
        self._check(lib().jp_write_vti(self.h, str(path).encode()))

    def write_boxes(self, path):
        self._check(lib().jp_write_boxes(self.h, str(path).encode()))

    def derived(self):
        d = Derived()
        self._check(lib().jp_get_derived(self.h, ctypes.byref(d)))
        return d.as_dict()

    def bench(self, steps, warmup=20, graph=False):
        r = BenchResult()
        self._check(lib().jp_bench(self.h, steps, warmup, 1 if graph else 0, ctypes.byref(r)))
        return r.as_dict()

    def population_bytes(self, level=0):
        n = ctypes.c_size_t()
        self._check(lib().jp_population_bytes(self.h, level, ctypes.byref(n)))
        return n.value

    def raw(self, level=0):
        nb = self.population_bytes(level)
        dt = RAW_DTYPE[self.prec]
        out = np.empty(nb // np.dtype(dt).itemsize, dtype=dt)
        if nb:
            self._check(lib().jp_get_populations(self.h, level, out.ctypes.data_as(ctypes.c_void_p), nb))
        return out

    def set_raw(self, a, level=0):
        nb = self.population_bytes(level)
        a = np.ascontiguousarray(a, dtype=RAW_DTYPE[self.prec])
        self._check(lib().jp_set_populations(self.h, level, a.ctypes.data_as(ctypes.c_void_p), a.nbytes)
                    if a.nbytes == nb else JP_ERR_ARG)

    def use_graph(self, on=True):
        self._check(lib().jp_use_graph(self.h, 1 if on else 0))

    def set_variant(self, v):
        self._check(lib().jp_set_variant(self.h, v))

    def set_step_pairs(self, on=True):
        """Two fused steps per launch (default where a column shape applies) or single steps."""
        self._check(lib().jp_set_step_pairs(self.h, 1 if on else 0))

    def set_step_pair_shape(self, pick):
        """Fused pairs with column shape `pick` of the library's list (0: 128 x 4, 1: 64 x 8, 2: 32 x 16 with the
        longest z-segment that divides N; 3-5: the same with half that segment; -1: the first that applies)."""
        self._check(lib().jp_set_step_pair_shape(self.h, pick))

    def step_pair_shape(self):
        """(x, y, z extent, lanes) of the fused pair's column; zeros when steps run singly."""
        a = (ctypes.c_int * 4)()
        self._check(lib().jp_step_pair_shape(self.h, ctypes.cast(a, ctypes.c_void_p)))
        return tuple(a)

    def steps_per_launch(self):
        """1 for single steps, 2 for fused pairs."""
        k = ctypes.c_int()
        self._check(lib().jp_steps_per_launch(self.h, ctypes.byref(k)))
        return k.value

    def sync(self):
        self._check(lib().jp_sync(self.h))


class Uniform(_Ctx):
    """Uniform periodic D3Q19 TGV lattice on the GPU (include/jp.h jp_uniform_create)."""

    def __init__(self, n, re=1600.0, u0=0.05, prec=PREC_FP32, phase=(0.0, 0.0, 0.0), scale=0.0, device=0,
                 plane_pad=0):
        self.prec, self.n = prec, n
        p = UniformParams(n=n, re=re, u0=u0, prec=prec, scale=scale, device=device, plane_pad=plane_pad)
        for k in range(3):
            p.phase[k] = phase[k]
        h = ctypes.c_void_p()
        rc = lib().jp_uniform_create(ctypes.byref(p), ctypes.byref(h))
        self.h = h.value
        self._check(rc)

    def autotune(self):
        """Times single steps in each launch variant and every fused column shape that divides N on the
        current state (which it leaves unchanged) and keeps the fastest; returns its shape (x, y, z extent,
        lanes), zeros for single steps. self.tuned then also holds the single-step variant and the waves per
        CU of the kept kernel."""
        g = (ctypes.c_int * ENS_KERNELS)()
        self._check(lib().jp_autotune(self.h, g))
        self.tuned = {"shape": tuple(g[:4]), "variant": g[4], "waves_per_cu": g[5], "steps_per_launch": g[6]}
        return tuple(g[:4])


def _fptr(a):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(ctypes.c_void_p)


class Amr(_Ctx):
    """3-level block AMR TGV on the GPU (include/jp.h jp_amr_create)."""

    def __init__(self, n=32, re=800.0, u0=0.05, l1=MODE_ALL, l2=MODE_ALL, frac1=0.35, frac2=0.35, adapt=100,
                 maxgb=0.0, prec=PREC_FP32, skipprep=0, skiprestr=0, phase=(0.0, 0.0, 0.0), scales=(0, 0, 0),
                 device=0):
        self.prec, self.n = prec, n
        p = AmrParams(n=n, re=re, u0=u0, prec=prec, l1_mode=l1, l2_mode=l2, frac1=frac1, frac2=frac2, adapt=adapt,
                      maxgb=maxgb, skipprep=skipprep, skiprestr=skiprestr, device=device)
        for k in range(3):
            p.phase[k] = phase[k]
            p.scale[k] = scales[k]
        h = ctypes.c_void_p()
        rc = lib().jp_amr_create(ctypes.byref(p), ctypes.byref(h))
        self.h = h.value
        self._check(rc)

    def init(self, sen0=None, sen1_by_pos=None):
        a0, a1 = _fptr(sen0), _fptr(sen1_by_pos)
        self._check(lib().jp_init_with_sensor(self.h, a0[1] if a0 else None, a1[1] if a1 else None))

    def regrid(self, sen0=None, sen1_by_pos=None):
        a0, a1 = _fptr(sen0), _fptr(sen1_by_pos)
        self._check(lib().jp_regrid_with_sensor(self.h, a0[1] if a0 else None, a1[1] if a1 else None))

    def counts(self):
        c = Counts()
        self._check(lib().jp_get_counts(self.h, ctypes.byref(c)))
        return c.as_dict()

    def table(self, name):
        which = TABLES.index(name)
        n = ctypes.c_long()
        dt = np.float32 if name.startswith("sen") else np.int32
        cap = 1 << 24
        buf = np.empty(cap, dtype=dt)
        self._check(lib().jp_get_table(self.h, which, buf.ctypes.data_as(ctypes.c_void_p), cap, ctypes.byref(n)))
        return buf[:n.value].copy()


class Member(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("re", ctypes.c_double), ("u0", ctypes.c_double), ("phase", ctypes.c_double * 3),
                ("amr", ctypes.c_int), ("l1_mode", ctypes.c_int), ("l2_mode", ctypes.c_int), ("frac1", ctypes.c_double),
                ("frac2", ctypes.c_double), ("adapt", ctypes.c_int)]


class _Borrowed(_Ctx):
    """A member of an ensemble: owned by the ensemble, never destroyed through this handle."""

    def __init__(self, h, prec, owner):
        self.h, self.prec, self._owner = h, prec, owner

    def close(self):
        self.h = None

    def counts(self):
        return Amr.counts(self)

    def table(self, name):
        return Amr.table(self, name)


class Ensemble(_Ctx):
    """Batched ensemble of uniform and AMR TGV lattices (include/jp.h jp_ensemble_create)."""

    def __init__(self, members, prec=PREC_FP32, maxgb=0.0, device=0):
        self.prec = prec
        arr = (Member * len(members))()
        for i, m in enumerate(members):
            arr[i].n = m["n"]
            arr[i].re = m["re"]
            arr[i].u0 = m["u0"]
            for k in range(3):
                arr[i].phase[k] = m.get("phase", (0.0, 0.0, 0.0))[k]
            arr[i].amr = 1 if m.get("amr") else 0
            arr[i].l1_mode = m.get("l1", MODE_SENSOR)
            arr[i].l2_mode = m.get("l2", MODE_SENSOR)
            arr[i].frac1 = m.get("frac1", 0.3)
            arr[i].frac2 = m.get("frac2", 0.1)
            arr[i].adapt = m.get("adapt", 10)
        h = ctypes.c_void_p()
        rc = lib().jp_ensemble_create(arr, len(members), prec, maxgb, device, ctypes.byref(h))
        self.h = h.value
        self._check(rc)
        self.members = []
        for i in range(len(members)):
            mh = ctypes.c_void_p()
            self._check(lib().jp_ensemble_member(self.h, i, ctypes.byref(mh)))
            self.members.append(_Borrowed(mh.value, prec, self))

    def close(self):
        for m in getattr(self, "members", []):
            m.close()
        super().close()

    def regrid(self):
        self._check(lib().jp_regrid(self.h))

    def use_graph(self, on=True):
        self._check(lib().jp_use_graph(self.h, 1 if on else 0))

    def autotune(self):
        g = (ctypes.c_int * ENS_KERNELS)()
        self._check(lib().jp_autotune(self.h, g))
        return list(g)

    def dense_shape(self, cells_per_lane=0):
        """Sets (1 or 8) or with 0 only reads the batched dense step's cells per lane; returns the shape in
        effect."""
        cur = ctypes.c_int()
        self._check(lib().jp_ensemble_dense_shape(self.h, cells_per_lane, ctypes.byref(cur)))
        return cur.value
