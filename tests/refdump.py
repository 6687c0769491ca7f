"""Readers for the binary dumps written by oracle/ref_harness2d.f90 and
oracle/ref_harness1d.f90 (layouts documented in those files)."""
import numpy as np


def read2d(path):
    with open(path, "rb") as f:
        nx, ny, nv, steps, it = np.frombuffer(f.read(20), dtype=np.int32)
        sim, dt, *res = np.frombuffer(f.read(48), dtype=np.float64)
        dts = np.frombuffer(f.read(8 * steps), dtype=np.float64)
        u = np.frombuffer(f.read(8 * nv), dtype=np.float64)
        v = np.frombuffer(f.read(8 * nv), dtype=np.float64)
        wt = np.frombuffer(f.read(8 * nv * nv), dtype=np.float64)
        ncell = nx * ny
        w = np.frombuffer(f.read(8 * 4 * ncell), dtype=np.float64).reshape(ncell, 4)
        h = np.frombuffer(f.read(8 * nv * nv * ncell), dtype=np.float64).reshape(ncell, nv * nv)
        b = np.frombuffer(f.read(8 * nv * nv * ncell), dtype=np.float64).reshape(ncell, nv * nv)
    return dict(nx=int(nx), ny=int(ny), nv=int(nv), steps=int(steps), iter=int(it),
                simtime=sim, dt=dt, res=np.array(res), dts=dts, u=u, v=v, weight=wt,
                w=w, h=h, b=b)


def read1d(path):
    with open(path, "rb") as f:
        nx, nv, steps, it = np.frombuffer(f.read(16), dtype=np.int32)
        vals = np.frombuffer(f.read(64), dtype=np.float64)
        dts = np.frombuffer(f.read(8 * steps), dtype=np.float64)
        u = np.frombuffer(f.read(8 * nv), dtype=np.float64)
        wt = np.frombuffer(f.read(8 * nv), dtype=np.float64)
        w = np.frombuffer(f.read(8 * 3 * nx), dtype=np.float64).reshape(nx, 3)
        h = np.frombuffer(f.read(8 * nv * nx), dtype=np.float64).reshape(nx, nv)
        b = np.frombuffer(f.read(8 * nv * nx), dtype=np.float64).reshape(nx, nv)
    return dict(nx=int(nx), nv=int(nv), steps=int(steps), iter=int(it), simtime=vals[0],
                dt=vals[1], res=vals[2:5], steady_res=vals[5:8], dts=dts, u=u, weight=wt,
                w=w, h=h, b=b)
