"""Writes the golden fixtures of tests/golden/*.npz by running the compiled
reference solver (oracle/_ref, built from the reference sources by
oracle/Makefile) on the namelists in tests/golden/cases/.

Each fixture holds the inputs (the namelist text, step count, velocity
grid) and the reference's outputs after that many steps: conserved
variables w, distributions h and b, the dt of every step, simtime, the
final iteration index and residual.

Fixtures of the named full-size configurations (LARGE) keep only the
conserved variables, the step sizes and the loop state (their h and b
would take hundreds of MB); they go to tests/golden/large/<case>.npz.

It also writes tests/golden/macro/<case>.npz: the macro<iter>.dat files
the reference's own main programs (oracle/_ref/ugks2d_ref, ugks1d_ref)
write for the namelists in MACRO_CASES, as text, for the driver tests.

Run in a container that has the reference sources:
    make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
import refdump  # noqa: E402

# name: (namelist, dim, steps, cartesian RT, seed)
FIXTURES = {
    "ref2d_poiseuille_smoke": ("ref2d_poiseuille_smoke.namelist", 2, 10, 0, 0),
    "ref2d_poiseuille_symb_smoke": ("ref2d_poiseuille_symb_smoke.namelist", 2, 10, 0, 0),
    "ref2d_rt_smoke": ("ref2d_rt_smoke.namelist", 2, 10, 0, 0),
    "p2d_poiseuille_4x8": ("p2d_poiseuille_4x8.namelist", 2, 25, 0, 0),
    "p2d_rt_cartesian_6x10": ("p2d_rt_cartesian_6x10.namelist", 2, 15, 1, 5),
    # the headline velocity grid and Hermite order: GH40, Nm = 12
    "p2d_rt_gh40_nm12_6x8": ("p2d_rt_gh40_nm12_6x8.namelist", 2, 6, 1, 3),
    # Newton-Cotes velocity grids (2d/src/velocity.f90:29-61, 1d:34-63)
    "p2d_poiseuille_nc21_4x8": ("p2d_poiseuille_nc21_4x8.namelist", 2, 20, 0, 0),
    "p2d_rt_cartesian_nc17_6x10": ("p2d_rt_cartesian_nc17_6x10.namelist", 2, 12, 1, 9),
    "p1d_sod_nc100_40": ("p1d_sod_nc100_40.namelist", 1, 50, 0, 0),
    # the finite-difference force (2d/src/gravity.f90:108-187, 1d:127-154)
    "p2d_poiseuille_fd_nc21_4x8": ("p2d_poiseuille_fd_nc21_4x8.namelist", 2, 20, 0, 0),
    "p2d_rt_cartesian_fd_gh12_6x10": ("p2d_rt_cartesian_fd_gh12_6x10.namelist", 2, 12, 1, 4),
    "p1d_sod_fd_nc100_40": ("p1d_sod_fd_nc100_40.namelist", 1, 50, 0, 0),
    "ref1d_fourier_smoke": ("ref1d_fourier_smoke.namelist", 1, 40, 0, 0),
    "p1d_sod_40": ("p1d_sod_40.namelist", 1, 60, 0, 0),
    "p1d_hydrostatic_30": ("p1d_hydrostatic_30.namelist", 1, 40, 0, 0),
}

# W-only fixtures of the named configurations at full size:
#   Poiseuille 4x64 GH24 Nm8, all 2000 steps;
#   Cartesian RT 120x120 GH28 Nm10, 12 steps;
#   a 512x16-cell band of the 512x512 RT (same cell size, GH40, Nm12), 4 steps
LARGE = {
    "c2_poiseuille_4x64": ("c2_poiseuille_4x64.namelist", 2, 2000, 0, 0),
    "c3_rt_120x120": ("c3_rt_120x120.namelist", 2, 12, 1, 1),
    "c5_band_512x16": ("c5_band_512x16.namelist", 2, 4, 1, 1),
}


# name: dim; the reference main program runs each namelist to its own end
# (the reference main program does not finish on the hermite_symmetric_b
# smoke namelist in this build, so that case has no macro fixture)
MACRO_CASES = {
    "ref2d_rt_smoke": 2,
    "ref2d_poiseuille_smoke": 2,
    "ref1d_fourier_smoke": 1,
    "p1d_sod_40": 1,
    "c1_sod_200": 1,
}


def run_macro(name):
    dim = MACRO_CASES[name]
    path = os.path.join(HERE, "cases", name + ".namelist")
    exe = os.path.join(ROOT, "oracle", "_ref", f"ugks{dim}d_ref")
    with tempfile.TemporaryDirectory() as td:
        r = subprocess.run([exe, path], capture_output=True, text=True, cwd=td, check=True,
                           timeout=120)
        files = sorted(f for f in os.listdir(td) if f.startswith("macro") and f.endswith(".dat"))
        out = {f: np.array(open(os.path.join(td, f)).read()) for f in files}
    os.makedirs(os.path.join(HERE, "macro"), exist_ok=True)
    np.savez_compressed(os.path.join(HERE, "macro", name + ".npz"), **out)
    return files, r.stdout.strip().splitlines()[-1:]


def run(name):
    large = name in LARGE
    nml, dim, steps, cart, seed = LARGE[name] if large else FIXTURES[name]
    path = os.path.join(HERE, "cases", nml)
    exe = os.path.join(ROOT, "oracle", "_ref", f"ref_harness{dim}d")
    with tempfile.TemporaryDirectory() as td:
        dump = os.path.join(td, "dump.bin")
        args = [exe, path, dump, str(steps)] + ([str(cart), str(seed)] if dim == 2 else [])
        r = subprocess.run(args, capture_output=True, text=True, cwd=td, check=True)
        d = refdump.read2d(dump) if dim == 2 else refdump.read1d(dump)
    out = {k: np.asarray(v) for k, v in d.items() if not (large and k in ("h", "b"))}
    out["namelist"] = np.array(open(path).read())
    out["dim"] = np.array(dim)
    out["rt_cartesian"] = np.array(cart)
    out["seed"] = np.array(seed)
    out["requested_steps"] = np.array(steps)
    dest = os.path.join(HERE, "large") if large else HERE
    os.makedirs(dest, exist_ok=True)
    np.savez_compressed(os.path.join(dest, name + ".npz"), **out)
    return r.stdout.strip().splitlines()[-2:]


if __name__ == "__main__":
    # arguments: fixture names (FIXTURES or LARGE), or macro:<case> for a
    # macro-file fixture
    for n in (sys.argv[1:] or list(FIXTURES) + list(LARGE) + ["macro:" + m for m in MACRO_CASES]):
        print(n, run_macro(n[6:]) if n.startswith("macro:") else run(n))
