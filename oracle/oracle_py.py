"""ctypes binding of the C oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke()
and by the CPU-baseline leg of bench.py, to check results against.  The
product does not import it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
_dp = C.POINTER(C.c_double)
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "cport"], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        V = C.c_void_p
        for dim in ("2d", "1d"):
            getattr(lib, f"oracle{dim}_create").restype = V
            getattr(lib, f"oracle{dim}_create").argtypes = [C.c_void_p]
            getattr(lib, f"oracle{dim}_destroy").argtypes = [V]
            getattr(lib, f"oracle{dim}_step").argtypes = [V, C.c_void_p]
            getattr(lib, f"oracle{dim}_get").argtypes = [V, _dp, _dp, _dp]
            getattr(lib, f"oracle{dim}_set").argtypes = [V, _dp, _dp, _dp]
            getattr(lib, f"oracle{dim}_time").argtypes = [V, _dp, C.POINTER(C.c_int32), _dp]
            getattr(lib, f"oracle{dim}_heat_flux").argtypes = [V, _dp]
        lib.oracle2d_flux_interior.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double,
                                               C.c_double, C.c_double] + [_dp] * 15
        lib.oracle2d_hermite_force.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp,
                                               C.c_double, C.c_double, _dp, _dp, _dp, _dp]
        lib.oracle2d_flux_wall.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double,
                                           C.c_double, C.c_double, _dp, _dp, _dp, _dp,
                                           C.c_double, _dp, _dp, _dp]
        lib.oracle1d_hermite_force.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp]
        _lib = lib
    return _lib


def ptr(a):
    return None if a is None else a.ctypes.data_as(_dp)


class Oracle:
    """FP64 CPU restatement of the reference step for one case."""

    def __init__(self, case):
        self.lib = load()
        self.case = case
        self.dim = case.desc.dim
        self.pref = f"oracle{self.dim}d"
        d = case.bind()
        self._desc = d
        self.o = getattr(self.lib, self.pref + "_create")(C.addressof(d))
        if not self.o:
            raise ValueError("oracle rejected the descriptor")

    def __del__(self):
        if getattr(self, "o", None):
            getattr(self.lib, self.pref + "_destroy")(self.o)
            self.o = None

    def step(self, n=1):
        from hugks.capi import StepInfo
        info = StepInfo()
        for _ in range(n):
            getattr(self.lib, self.pref + "_step")(self.o, C.addressof(info))
        return info

    def get(self):
        c = self.case
        w = np.empty(c.ncell * (self.dim + 2))
        h = np.empty(c.ncell * c.nn)
        b = np.empty(c.ncell * c.nn)
        getattr(self.lib, self.pref + "_get")(self.o, ptr(w), ptr(h), ptr(b))
        return (w.reshape(c.ncell, self.dim + 2), h.reshape(c.ncell, c.nn),
                b.reshape(c.ncell, c.nn))

    def set(self, w=None, h=None, b=None):
        conv = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        w, h, b = conv(w), conv(h), conv(b)
        getattr(self.lib, self.pref + "_set")(self.o, ptr(w), ptr(h), ptr(b))

    def heat_flux(self):
        q = np.empty(self.case.ncell * self.dim)
        getattr(self.lib, self.pref + "_heat_flux")(self.o, ptr(q))
        return q.reshape(self.case.ncell, self.dim)

    def time(self):
        t, dt = C.c_double(), C.c_double()
        it = C.c_int32()
        getattr(self.lib, self.pref + "_time")(self.o, C.byref(t), C.byref(it), C.byref(dt))
        return t.value, it.value, dt.value
