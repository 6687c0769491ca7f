"""The Fortran drivers over the C-ABI (hermite-ugks_amd/driver/ugks2d_gpu,
ugks1d_gpu) against the output files of the reference's own main programs
on the same namelists (tests/golden/macro/<case>.npz, written by
oracle/_ref/ugks2d_ref and ugks1d_ref through tests/golden/make_golden.py):
the same macro<iter>.dat files, the same header lines and layout, and every
number within 1e-10 of its block's scale (FP64 path)."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
DRIVER = {2: os.path.join(ROOT, "hermite-ugks_amd", "driver", "ugks2d_gpu"),
          1: os.path.join(ROOT, "hermite-ugks_amd", "driver", "ugks1d_gpu")}

RUNS = [
    ("ref2d_rt_smoke", 2),
    ("ref2d_poiseuille_smoke", 2),
    ("ref1d_fourier_smoke", 1),
    ("p1d_sod_40", 1),
    ("c1_sod_200", 1),
]

NUM = re.compile(r"[-+]?\d\.\d+E[-+]\d+")


def parse_macro(text, dim):
    """Header lines, the value blocks (2D: X, Y at the (I x J) nodes, then
    11 cell-centred fields; 1D: X and 9 fields) and the token count per
    line."""
    lines = text.splitlines()
    zone = [int(x) for x in re.findall(r"[IJ] =\s*(\d+)", lines[1])]
    vals = np.array([float(x) for x in NUM.findall("\n".join(lines[2:]))])
    if dim == 2:
        nn, nc = zone[0] * zone[1], (zone[0] - 1) * (zone[1] - 1)
        sizes = [nn, nn] + [nc] * 11
    else:
        sizes = [zone[0]] * 10
    assert vals.size == sum(sizes), (vals.size, sizes)
    blocks = np.split(vals, np.cumsum(sizes)[:-1])
    return lines[:2], blocks, [len(ln.split()) for ln in lines[2:]]


@pytest.mark.parametrize("name,dim", RUNS, ids=[r[0] for r in RUNS])
def test_driver_matches_reference_program(have_gpu, tmp_path, name, dim):
    exe = DRIVER[dim]
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built")
    nml = os.path.join(GOLDEN, "cases", name + ".namelist")
    print(f"driver {name} dim={dim}", flush=True)
    r = subprocess.run([exe, nml], cwd=tmp_path, capture_output=True, text=True, timeout=90)
    assert r.returncode == 0, f"rc={r.returncode}\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}"
    ref = np.load(os.path.join(GOLDEN, "macro", name + ".npz"), allow_pickle=False)
    files = sorted(ref.files)
    assert files == sorted(p.name for p in tmp_path.glob("macro*.dat")), r.stdout[-2000:]
    for fname in files:
        h1, a, shape1 = parse_macro((tmp_path / fname).read_text(), dim)
        h2, b, shape2 = parse_macro(str(ref[fname]), dim)
        assert h1 == h2, fname
        assert shape1 == shape2, fname
        floor = 1e-13 * max(np.max(np.abs(b[2 if dim == 2 else 1])), 1.0)
        for k in range(len(b)):
            scale = np.max(np.abs(b[k]))
            err = np.max(np.abs(a[k] - b[k]))
            assert err <= 1e-10 * scale + floor, (fname, k, err, scale)


def test_driver_restart_continues_the_run(have_gpu, tmp_path):
    """&gpu checkpoint_file / restart_file: 6 steps, save, restart for 4
    more, writes the same final macro file as 10 steps in one run."""
    exe = DRIVER[2]
    base = open(os.path.join(GOLDEN, "cases", "ref2d_poiseuille_smoke.namelist")).read()
    base = re.sub(r"&gpu.*?/", "", base, flags=re.S).replace("maxtime=0.001", "maxtime=1.0")

    def run(sub, gpu):
        d = tmp_path / sub
        d.mkdir()
        (d / "in.namelist").write_text(base + "\n&gpu\n" + gpu + "\n/\n")
        r = subprocess.run([exe, "in.namelist"], cwd=d, capture_output=True, text=True, timeout=90)
        assert r.returncode == 0, f"rc={r.returncode}\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}"
        return d

    print("driver restart", flush=True)
    one = run("one", "max_steps = 10")
    first = run("first", "max_steps = 6, checkpoint_file = 'state.ckpt'")
    second = tmp_path / "second"
    second.mkdir()
    (second / "state.ckpt").write_bytes((first / "state.ckpt").read_bytes())
    (second / "in.namelist").write_text(
        base + "\n&gpu\nmax_steps = 4, restart_file = 'state.ckpt'\n/\n")
    r = subprocess.run([exe, "in.namelist"], cwd=second, capture_output=True, text=True, timeout=90)
    assert r.returncode == 0, r.stdout[-2000:]
    assert "restarted from" in r.stdout
    assert (first / "macro7.dat").exists() and (one / "macro11.dat").exists()
    assert (second / "macro11.dat").read_text() == (one / "macro11.dat").read_text()
