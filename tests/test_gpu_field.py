"""GPU parity of the frame output (the reference's --vtk path, src/tgv3d_uniform.cu:96-140 and
src/tgv3d_amr.cu:551-641): the output field computed on the device against the oracle from the same stored
state (Q bit-exact: the same FP64 operations in the same order; speed within one FP32 ulp, the FP64 square
root being the only library call; velocity exact), and the files read back the way the reference's
scripts read them (scripts/preview_tgv3d.py:21-42)."""
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import jpabi
import oracle as orc

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "jp-solver_amd", "build")


def read_vti(path):
    """scripts/preview_tgv3d.py:21-29 (WholeExtent, the '_' that opens the appended block, UInt64 byte
    counts), extended to every appended array by its offset."""
    raw = open(path, "rb").read()
    n = int(re.search(rb'WholeExtent="0 (\d+)', raw).group(1)) + 1
    us = raw.find(b"_", raw.find(b"<AppendedData")) + 1
    arrays = {}
    pat = rb'<DataArray type="Float32" Name="(\w+)"(?: NumberOfComponents="(\d+)")? format="appended" offset="(\d+)"/>'
    for m in re.finditer(pat, raw):
        off = int(m.group(3))
        (nb,) = struct.unpack_from("<Q", raw, us + off)
        arrays[m.group(1).decode()] = (int(m.group(2) or 1), np.frombuffer(raw, "<f4", nb // 4, us + off + 8))
    return n, arrays, raw


def read_boxes(path):
    """scripts/preview_tgv3d.py:31-42: 8 ascii points per box."""
    txt = open(path).read()
    pts = np.array(re.search(r'<Points>.*?format="ascii">(.*?)</DataArray>', txt, re.S).group(1).split(), float)
    pts = pts.reshape(-1, 8, 3)
    lv = np.array(re.search(r'Name="level" format="ascii">(.*?)</DataArray>', txt, re.S).group(1).split(), int)
    return pts, lv.reshape(-1, 8)


def ulps(a, b):
    return np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64))


@pytest.mark.parametrize("prec", [jpabi.PREC_FP32, jpabi.PREC_FP16])
def test_uniform_field_matches_oracle(prec):
    o = orc.Uniform(24, prec=prec)
    o.init()
    o.step(5)
    g = jpabi.Uniform(24, prec=prec)
    g.set_raw(o.raw())
    gf, of = g.field(velocity=True), o.field()
    assert np.array_equal(gf["Q"], of["Q"])
    assert np.array_equal(gf["velocity"], of["velocity"])
    assert ulps(gf["speed"], of["speed"]).max() <= 1


def test_amr_field_matches_oracle():
    kw = dict(n=16, re=800.0, l1=jpabi.MODE_SLAB, l2=jpabi.MODE_ALL)
    g, o = jpabi.Amr(maxgb=8.0, **kw), orc.Amr(budget_bytes=8.0 * 2 ** 30, **kw)
    g.init()
    o.init()
    o.step(2)
    for lvl in range(3):
        g.set_raw(o.raw(lvl), lvl)
    gf, of = g.field(), o.field()
    assert gf["Q"].shape == (64, 64, 64)
    assert np.array_equal(gf["Q"], of["Q"])
    assert ulps(gf["speed"], of["speed"]).max() <= 1


def test_vti_frames_read_back(tmp_path):
    u = jpabi.Uniform(16)
    u.init()
    u.step(3)
    p = tmp_path / "u.vti"
    u.write_vti(p)
    n, arrays, raw = read_vti(p)
    f = u.field(velocity=True)
    assert n == 16 and list(arrays) == ["Q", "speed", "velocity"]
    assert b'header_type="UInt64"' in raw and b'<PointData Scalars="Q" Vectors="velocity">' in raw
    assert np.array_equal(arrays["Q"][1], f["Q"].ravel()) and np.array_equal(arrays["speed"][1], f["speed"].ravel())
    assert arrays["velocity"][0] == 3 and np.array_equal(arrays["velocity"][1], f["velocity"].ravel())
    a = jpabi.Amr(8, re=800.0, l1=jpabi.MODE_SLAB, l2=jpabi.MODE_ALL, maxgb=8.0)
    a.init()
    p = tmp_path / "a.vti"
    a.write_vti(p)
    n, arrays, raw = read_vti(p)
    f = a.field()
    assert n == 32 and list(arrays) == ["Q", "speed"] and b'<PointData Scalars="Q">' in raw
    assert np.array_equal(arrays["Q"][1], f["Q"].ravel()) and np.array_equal(arrays["speed"][1], f["speed"].ravel())


def test_level_boxes_follow_the_tables(tmp_path):
    n0 = 32
    rng = np.random.default_rng(3)
    a = jpabi.Amr(n0, re=1600.0, l1=jpabi.MODE_SENSOR, l2=jpabi.MODE_SENSOR, frac1=0.3, frac2=0.3, maxgb=8.0)
    a.init(rng.random((n0 // 8) ** 3).astype(np.float32), rng.random((n0 // 4) ** 3).astype(np.float32))
    c = a.counts()
    p = tmp_path / "b.vtp"
    a.write_boxes(p)
    pts, lv = read_boxes(p)
    assert pts.shape[0] == c["nf1"] + c["nf2"] and np.all(lv == lv[:, :1])
    lo, size = pts.min(axis=1), pts.max(axis=1) - pts.min(axis=1)
    cgx, l1gx = n0 // 8, n0 // 4
    pcb, oct1, f1pos = a.table("l1pcb"), a.table("l1oct"), a.table("f1pos")
    for b in range(c["nf1"]):  # level-1 tile: 16 finest cells, origin 2 x its level-1 coordinate
        cb, o = pcb[b], oct1[b]
        want = [2 * ((cb % cgx) * 16 + (o & 1) * 8), 2 * (((cb // cgx) % cgx) * 16 + ((o >> 1) & 1) * 8),
                2 * ((cb // cgx ** 2) * 16 + (o >> 2) * 8)]
        assert lv[b, 0] == 1 and np.array_equal(lo[b], want) and np.all(size[b] == 16)
    pl1, oct2 = a.table("l2pl1"), a.table("l2oct")
    for k in range(c["nf2"]):
        b, o = c["nf1"] + k, oct2[k]
        p1 = f1pos[pl1[k]]
        want = [(p1 % l1gx) * 16 + (o & 1) * 8, ((p1 // l1gx) % l1gx) * 16 + ((o >> 1) & 1) * 8,
                (p1 // l1gx ** 2) * 16 + (o >> 2) * 8]
        assert lv[b, 0] == 2 and np.array_equal(lo[b], want) and np.all(size[b] == 8)


def test_drivers_write_frames(tmp_path):
    # src/tgv3d_uniform.cu:165, 193-196 and src/tgv3d_amr.cu:652, 869, 889-895: frame every steps/frames steps
    r = subprocess.run([os.path.join(BUILD, "uniform"), "--n", "16", "--tstar", "0.3", "--vtk", str(tmp_path / "u"),
                        "--frames", "3"], capture_output=True, text=True, timeout=120)
    assert "wrote 3 vti frames" in r.stdout, r.stdout + r.stderr
    assert sorted(os.listdir(tmp_path)) == ["u_0000.vti", "u_0001.vti", "u_0002.vti"]
    r = subprocess.run([os.path.join(BUILD, "amr"), "--n", "8", "--tstar", "0.3", "--l1", "slab", "--vtk",
                        str(tmp_path / "a"), "--frames", "2"], capture_output=True, text=True, timeout=120)
    assert "wrote 2 frames" in r.stdout, r.stdout + r.stderr
    names = sorted(x for x in os.listdir(tmp_path) if x.startswith("a_"))
    assert names == ["a_0000.vti", "a_0000.vtp", "a_0001.vti", "a_0001.vtp"]
    n, arrays, _ = read_vti(tmp_path / "a_0001.vti")
    assert n == 32 and np.all(np.isfinite(arrays["Q"][1]))
