"""The Fortran drivers over the C-ABI (hermite-ugks_amd/driver/ugks2d_gpu,
ugks1d_gpu) against the reference's own main programs built from its
sources (oracle/_ref/ugks2d_ref, ugks1d_ref) on the same namelist: the
same macro<iter>.dat files, the same header lines and layout, and every
number within 1e-10 of the block's scale (FP64 path)."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CASES = os.path.join(ROOT, "tests", "golden", "cases")
DRIVER = {2: os.path.join(ROOT, "hermite-ugks_amd", "driver", "ugks2d_gpu"),
          1: os.path.join(ROOT, "hermite-ugks_amd", "driver", "ugks1d_gpu")}
REFPROG = {2: os.path.join(ROOT, "oracle", "_ref", "ugks2d_ref"),
           1: os.path.join(ROOT, "oracle", "_ref", "ugks1d_ref")}

RUNS = [
    ("ref2d_rt_smoke", 2),
    ("ref2d_poiseuille_smoke", 2),
    ("ref2d_poiseuille_symb_smoke", 2),
    ("ref1d_fourier_smoke", 1),
    ("p1d_sod_40", 1),
    ("c1_sod_200", 1),
]

NUM = re.compile(r"[-+]?\d\.\d+E[-+]\d+")


def run(exe, nml, cwd):
    r = subprocess.run([exe, nml], cwd=cwd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, f"{exe} rc={r.returncode}\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}"
    return r.stdout


def read_macro(path, nblock):
    with open(path) as f:
        lines = f.read().splitlines()
    vals = np.array([float(x) for x in NUM.findall("\n".join(lines[2:]))])
    return lines[:2], vals.reshape(nblock, -1), [len(ln.split()) for ln in lines[2:]]


@pytest.mark.parametrize("name,dim", RUNS, ids=[r[0] for r in RUNS])
def test_driver_matches_reference_program(have_gpu, tmp_path, name, dim):
    for exe in (DRIVER[dim], REFPROG[dim]):
        if not os.path.exists(exe):
            pytest.fail(f"{exe} is not built")
    nml = os.path.join(CASES, name + ".namelist")
    ours, theirs = tmp_path / "ours", tmp_path / "ref"
    ours.mkdir()
    theirs.mkdir()
    print(f"driver {name} dim={dim}", flush=True)
    out = run(DRIVER[dim], nml, ours)
    run(REFPROG[dim], nml, theirs)
    files = sorted(p.name for p in theirs.glob("macro*.dat"))
    assert files and files == sorted(p.name for p in ours.glob("macro*.dat")), out[-2000:]
    nblock = 13 if dim == 2 else 10
    for fname in files:
        h1, a, shape1 = read_macro(ours / fname, nblock)
        h2, b, shape2 = read_macro(theirs / fname, nblock)
        assert h1 == h2, fname
        assert shape1 == shape2, fname
        floor = 1e-13 * max(np.max(np.abs(b[2 if dim == 2 else 1])), 1.0)
        for k in range(nblock):
            scale = np.max(np.abs(b[k]))
            err = np.max(np.abs(a[k] - b[k]))
            assert err <= 1e-10 * scale + floor, (fname, k, err, scale)
