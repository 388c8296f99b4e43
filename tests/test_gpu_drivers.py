"""The ./uniform and ./amr drivers keep the reference's stdout contract: every line they print matches one
of the reference's format strings (src/tgv3d_uniform.cu:154-155, 188, 197-201; src/tgv3d_amr.cu:818-819,
861-862, 887, 898-905), with the defaults of src/tgv3d_amr.cu:644-656 (--frames 240) and
src/tgv3d_uniform.cu:143-147 (--frames 160)."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BUILD = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "jp-solver_amd", "build")
F = r"-?\d+\.\d+"          # %.Nf
E = r"-?\d\.\d+e[+-]\d+"   # %.Ne
G = r"[-+0-9.e]+"          # %g

UNIFORM = [
    rf"\[jp-solver uniform\] TGV3D N=\d+ MB=8 blocks=\d+ Re={G} nu={G} tau={F} steps=\d+ t\*max={F}(  \(writing vti\))?",
    rf"  t\*={F}  E/E0={F}  enstrophy={E}  mass_drift={E}",
    rf"  enstrophy peak at t\*={F} \(value {E}\)",
    rf"  energy monotonic: (yes|NO)   mass drift {E}",
    r"  (VALID \(enstrophy peaks then falls, energy monotonic, mass held in FP32\)|CHECK \(see values\))",
    r"  wrote \d+ vti frames \S+_0000\.vti \.\.",
]
AMR = [
    rf"\[jp-solver amr\] 3-level 3D\. N0=\d+ N1=\d+ N2=\d+ Re={G} L1=\w+\(\d+ blk\) L2=\w+\(\d+ blk\) tau0={F} "
    r"steps=\d+( \(vti\+vtp\))?",
    rf"  t\*={F} E/E0={F} enstrophy={E} mass_drift={E}",
    rf"  enstrophy peak at t\*={F} \(value {E}\)",
    rf"  energy monotonic: (yes|NO)  mass drift {E}",
    r"  pool high-water: L1 \d+/\d+ blocks \(\d+%\), L2 \d+/\d+ \(\d+%\)  \[full-refine would be \d+/\d+\]",
    r"  (VALID \(breakdown present, energy monotonic, mass held\)|CHECK)",
    r"  wrote \d+ frames \S+_0000\.vti/\.vtp \.\.",
]
AMR_BENCH = [
    AMR[0],
    rf"\[bench(\+graph)?\] AMR N0=\d+\(fine \d+\): {F} s  work={E} updates  {F} MLUPS\(work\)  pools {F} GB  "
    rf"ctx\+pools {F} GB  \(L1=\d+ L2=\d+\)",
]
EXTRA = rf"  {F} GB/s at \d+ B/update \({F}% of 8 TB/s HBM\)"  # the build's roofline line after the bench line


def run(args):
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode in (0, 1), r.stdout + r.stderr
    return r.stdout.splitlines()


def check(lines, pats):
    for ln in lines:
        assert any(re.fullmatch(p, ln) for p in pats), f"line not in the reference's formats: {ln!r}"


def test_uniform_stdout_and_frames_default(tmp_path):
    lines = run([os.path.join(BUILD, "uniform"), "--n", "16", "--tstar", "0.4", "--vtk", str(tmp_path / "u")])
    check(lines, UNIFORM)
    steps = int(re.search(r"steps=(\d+)", lines[0]).group(1))
    frame_k = max(1, steps // 160)
    want = min(160, steps // frame_k + 1)
    assert lines[-1] == f"  wrote {want} vti frames {tmp_path / 'u'}_0000.vti .."
    assert lines[0].endswith("  (writing vti)")


def test_amr_stdout_and_frames_default(tmp_path):
    lines = run([os.path.join(BUILD, "amr"), "--n", "16", "--tstar", "0.4", "--l1", "slab", "--vtk", str(tmp_path / "a")])
    check(lines, AMR)
    assert lines[0].endswith(" (vti+vtp)")
    steps = int(re.search(r"steps=(\d+)", lines[0]).group(1))
    frame_k = max(1, steps // 240)
    want = min(240, steps // frame_k + 1)  # src/tgv3d_amr.cu:652, 869
    assert lines[-1] == f"  wrote {want} frames {tmp_path / 'a'}_0000.vti/.vtp .."
    lines = run([os.path.join(BUILD, "amr"), "--n", "16", "--tstar", "0.2"])
    check(lines, AMR)
    assert not lines[0].endswith(")")


@pytest.mark.parametrize("graph", ["0", "1"])
def test_amr_bench_line(graph):
    lines = run([os.path.join(BUILD, "amr"), "--n", "24", "--re", "1600", "--l1", "sensor", "--l2", "sensor", "--frac1", "0.3",
                 "--frac2", "0.1", "--adapt", "10", "--bench", "1", "--bench-steps", "40", "--graph", graph])
    assert len(lines) == 3
    check(lines[:2], AMR_BENCH)
    assert lines[1].startswith("[bench+graph]" if graph == "1" else "[bench]")
    assert re.fullmatch(EXTRA, lines[2]), lines[2]
    used = float(re.search(r"ctx\+pools (\S+) GB", lines[1]).group(1))
    pools = float(re.search(r" pools (\S+) GB", lines[1]).group(1))
    assert used >= pools > 0
