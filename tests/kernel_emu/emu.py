"""Runs a case through the CPU-emulation build of the library
(tests/kernel_emu/_build/emu_driver).  TEST INFRASTRUCTURE ONLY."""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "_build", "emu_driver")


def build():
    subprocess.run([os.path.join(HERE, "build.sh")], check=True, capture_output=True)


def run(case, steps, timeout=600):
    if not os.path.exists(DRIVER):
        build()
    d = case.desc
    a = case.arrays
    with tempfile.TemporaryDirectory() as td:
        cin, cout = os.path.join(td, "case.bin"), os.path.join(td, "out.bin")
        with open(cin, "wb") as f:
            f.write(bytes(d))
            f.write(np.asarray(a["u"], np.float64).tobytes())
            if d.dim == 2:
                f.write(np.asarray(a["v"], np.float64).tobytes())
            f.write(np.asarray(a["weight"], np.float64).tobytes())
            f.write(np.asarray(a["prim0"], np.float64).tobytes())
            if d.dim == 2:
                for k in ("accel_cell", "accel_vface", "accel_hface"):
                    f.write(np.asarray(a[k], np.float64).tobytes())
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0")
        r = subprocess.run([DRIVER, cin, str(steps), cout], capture_output=True, text=True,
                           timeout=timeout, env=env)
        if not os.path.exists(cout):
            raise RuntimeError(f"emu_driver rc={r.returncode}\n{r.stderr[-4000:]}")
        ncell, nn, dim = case.ncell, case.nn, d.dim
        with open(cout, "rb") as f:
            rc, st = np.frombuffer(f.read(8), np.int32)
            simtime, dt = np.frombuffer(f.read(16), np.float64)
            w = np.frombuffer(f.read(8 * (dim + 2) * ncell), np.float64).reshape(ncell, dim + 2)
            q = np.frombuffer(f.read(8 * dim * ncell), np.float64).reshape(ncell, dim)
            h = np.frombuffer(f.read(8 * ncell * nn), np.float64).reshape(ncell, nn)
            b = np.frombuffer(f.read(8 * ncell * nn), np.float64).reshape(ncell, nn)
        return dict(rc=int(rc), steps=int(st), simtime=simtime, dt=dt, w=w, q=q, h=h, b=b,
                    stderr=r.stderr, returncode=r.returncode)
