"""ctypes binding of libhugks.so (include/hugks.h).

The structures mirror the C declarations field for field; ``Solver`` wraps
one handle.  Loading fails loudly when the library is missing: there is no
fallback implementation.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.environ.get("HUGKS_LIB") or os.path.join(os.path.dirname(HERE), "libhugks.so")

ABI_VERSION = 1
OK = 0
E_ARG, E_NOMEM, E_DEVICE, E_UNSUPPORTED, E_NUMERIC, E_WALL, E_NOHIP = -1, -2, -3, -4, -5, -6, -7
E_IO = -8
FP64, MIXED_FP32, MIXED_BF16X3 = 0, 1, 2
BC_PERIODIC, BC_SPECULAR, BC_DIFFUSE = 0, 1, 2
TAU_MU_OVER_RHO, TAU_VHS = 0, 1
DT_2D_GENERAL, DT_2D_WALL_NORMAL, DT_1D = 0, 1, 2
FORCE_HERMITE_SYMMETRIC, FORCE_HERMITE_SYMMETRIC_B, FORCE_HERMITE_LEGACY = 0, 1, 2
FORCE_FINITE_DIFFERENCE = 3
STOP_NONE, STOP_MAXTIME, STOP_RESIDUAL, STOP_TIME, STOP_STEADY, STOP_ERROR = 0, 1, 2, 3, 4, 5
FLUX_FUSED, FLUX_SPLIT, FLUX_STENCIL, FLUX_CELLSUM = 0, 1, 2, 3
UPDATE_LDS, UPDATE_MFMA, UPDATE_MFMA_NODE, UPDATE_MFMA_FACE = 0, 1, 2, 3
BATCH_AUTOTUNE = 1

_dp = C.POINTER(C.c_double)


class Desc(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32), ("dim", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32),
        ("nv", C.c_int32), ("nm", C.c_int32), ("precision", C.c_int32),
        ("tau_model", C.c_int32), ("dt_rule", C.c_int32), ("clamp_dt", C.c_int32),
        ("stop_on_residual", C.c_int32), ("bc", C.c_int32 * 4), ("force_variant", C.c_int32),
        ("steady_steps", C.c_int32), ("reserved_i", C.c_int32 * 1),
        ("dx", C.c_double), ("dy", C.c_double), ("pi", C.c_double), ("smv", C.c_double),
        ("innerk", C.c_double), ("gamma", C.c_double), ("prandtl", C.c_double),
        ("shakhov_coef", C.c_double), ("muref", C.c_double), ("omega", C.c_double),
        ("cfl", C.c_double), ("maxtime", C.c_double), ("eps_res", C.c_double),
        ("umax_dt", C.c_double), ("vmax_dt", C.c_double), ("force_rate", C.c_double),
        ("tw", C.c_double * 4), ("tref", C.c_double), ("phi", C.c_double),
        ("steady_tol", C.c_double), ("mass_tol", C.c_double), ("fd_span", C.c_double),
        ("reserved_d", C.c_double * 1),
        ("u", _dp), ("v", _dp), ("weight", _dp), ("prim0", _dp), ("accel_cell", _dp),
        ("accel_vface", _dp), ("accel_hface", _dp),
    ]


class StepInfo(C.Structure):
    _fields_ = [
        ("steps", C.c_int32), ("iter", C.c_int32), ("stop_reason", C.c_int32),
        ("device_flags", C.c_int32), ("simtime", C.c_double), ("dt", C.c_double),
        ("res", C.c_double * 4), ("steady_res", C.c_double * 3), ("mass_rel_change", C.c_double),
    ]

    def as_dict(self):
        return {"steps": self.steps, "iter": self.iter, "stop_reason": self.stop_reason,
                "device_flags": self.device_flags, "simtime": self.simtime, "dt": self.dt,
                "res": list(self.res), "steady_res": list(self.steady_res),
                "mass_rel_change": self.mass_rel_change}


EXPORTS = [
    "hugks_create", "hugks_destroy", "hugks_last_error", "hugks_step", "hugks_advance",
    "hugks_advance_until",
    "hugks_get_macro", "hugks_get_macro_noalloc", "hugks_get_f", "hugks_set_f", "hugks_get_time",
    "hugks_checkpoint_write", "hugks_checkpoint_read", "hugks_kernel_times",
    "hugks_reset_kernel_times", "hugks_set_timing", "hugks_device_info", "hugks_debug_fetch",
    "hugks_set_schedule", "hugks_get_schedule", "hugks_set_update_waves",
    "hugks_get_update_waves", "hugks_tune", "hugks_tune_tiles", "hugks_batch_create",
    "hugks_batch_destroy", "hugks_batch_last_error", "hugks_batch_size", "hugks_batch_grouped",
    "hugks_batch_handle",
    "hugks_batch_advance", "hugks_batch_set_timing", "hugks_batch_kernel_times",
    "hugks_batch_reset_kernel_times", "hugks_write_blocks_async", "hugks_write_wait",
]
ARR_SHX, ARR_SHY, ARR_SBX, ARR_SBY, ARR_VFH, ARR_VFB, ARR_HFH, ARR_HFB, ARR_VFLUX, ARR_HFLUX = range(10)

_lib = None


def load(path=LIBPATH):
    """Loads libhugks.so once; raises if it is missing or incomplete."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libhugks.so not built ({path}); run __graft_entry__.build()")
    lib = C.CDLL(path)
    for name in EXPORTS:
        if not hasattr(lib, name):
            raise RuntimeError(f"libhugks.so lacks {name}")
    H = C.c_void_p
    lib.hugks_create.argtypes = [C.POINTER(Desc), C.POINTER(H)]
    lib.hugks_destroy.argtypes = [H]
    lib.hugks_last_error.argtypes = [H]
    lib.hugks_last_error.restype = C.c_char_p
    lib.hugks_step.argtypes = [H, C.POINTER(StepInfo)]
    lib.hugks_advance.argtypes = [H, C.c_int32, C.POINTER(StepInfo)]
    lib.hugks_advance_until.argtypes = [H, C.c_int32, C.c_double, C.POINTER(StepInfo)]
    lib.hugks_get_macro.argtypes = [H, _dp, _dp]
    lib.hugks_get_macro_noalloc.argtypes = [H, _dp, _dp]
    lib.hugks_get_f.argtypes = [H, _dp, _dp]
    lib.hugks_set_f.argtypes = [H, _dp, _dp, _dp]
    lib.hugks_get_time.argtypes = [H, _dp, C.POINTER(C.c_int32)]
    lib.hugks_checkpoint_write.argtypes = [H, C.c_char_p]
    lib.hugks_checkpoint_read.argtypes = [H, C.c_char_p]
    lib.hugks_kernel_times.argtypes = [H, C.c_int32, C.POINTER(C.c_char_p), _dp,
                                       C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    lib.hugks_reset_kernel_times.argtypes = [H]
    lib.hugks_set_timing.argtypes = [H, C.c_int32]
    lib.hugks_device_info.restype = C.c_char_p
    lib.hugks_debug_fetch.argtypes = [H, C.c_int32, _dp]
    I = C.POINTER(C.c_int32)
    lib.hugks_set_schedule.argtypes = [H, C.c_int32, C.c_int32]
    lib.hugks_get_schedule.argtypes = [H, I, I]
    lib.hugks_set_update_waves.argtypes = [H, C.c_int32]
    lib.hugks_get_update_waves.argtypes = [H, I]
    lib.hugks_tune.argtypes = [C.POINTER(Desc), I, I, _dp]
    lib.hugks_tune_tiles.argtypes = [C.POINTER(Desc), I, I, I, _dp]
    lib.hugks_batch_create.argtypes = [C.c_int32, C.POINTER(Desc), C.c_int32, C.POINTER(H)]
    lib.hugks_batch_destroy.argtypes = [H]
    lib.hugks_batch_last_error.argtypes = [H]
    lib.hugks_batch_last_error.restype = C.c_char_p
    lib.hugks_batch_size.argtypes = [H]
    lib.hugks_batch_size.restype = C.c_int32
    lib.hugks_batch_grouped.argtypes = [H]
    lib.hugks_batch_grouped.restype = C.c_int32
    lib.hugks_batch_handle.argtypes = [H, C.c_int32, C.POINTER(H)]
    lib.hugks_batch_advance.argtypes = [H, C.c_int32, C.POINTER(StepInfo)]
    lib.hugks_batch_set_timing.argtypes = [H, C.c_int32]
    lib.hugks_batch_kernel_times.argtypes = [H, C.c_int32, C.POINTER(C.c_char_p), _dp,
                                             C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    lib.hugks_batch_reset_kernel_times.argtypes = [H]
    lib.hugks_write_blocks_async.argtypes = [H, C.c_char_p, C.c_char_p, C.c_int32, C.c_int32,
                                             C.POINTER(C.c_int64), _dp]
    lib.hugks_write_wait.argtypes = [H]
    _lib = lib
    return lib


def write_blocks(path, header, per_line, blocks, lib=None, h=None):
    """hugks_write_blocks_async; with h None the file is written before
    the call returns.  Returns the status code."""
    lib = lib or load()
    vals = np.ascontiguousarray(np.concatenate([np.ravel(b) for b in blocks])
                                if blocks else np.zeros(0), dtype=np.float64)
    lens = (C.c_int64 * max(len(blocks), 1))(*[int(np.size(b)) for b in blocks])
    return lib.hugks_write_blocks_async(h, str(path).encode(), header.encode(), int(per_line),
                                        len(blocks), lens, _ptr(vals))


class HugksError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hugks error {code}: {msg}")
        self.code = code


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_dp)


class Case:
    """A fully specified case: descriptor scalars plus the arrays it points
    to (kept alive here)."""

    def __init__(self, desc, arrays, meta=None):
        self.desc = desc
        self.arrays = arrays
        self.meta = meta or {}

    @property
    def ncell(self):
        return self.desc.nx * self.desc.ny

    @property
    def nn(self):
        return self.desc.nv ** self.desc.dim

    def with_precision(self, precision):
        d = Desc.from_buffer_copy(self.desc)
        d.precision = precision
        return Case(d, self.arrays, dict(self.meta))

    def bind(self):
        """Points the descriptor at the numpy arrays and returns it."""
        a = self.arrays
        d = self.desc
        d.u = _ptr(a["u"])
        d.v = _ptr(a.get("v"))
        d.weight = _ptr(a["weight"])
        d.prim0 = _ptr(a["prim0"])
        d.accel_cell = _ptr(a.get("accel_cell"))
        d.accel_vface = _ptr(a.get("accel_vface"))
        d.accel_hface = _ptr(a.get("accel_hface"))
        return d


class Solver:
    """One library handle (one case on the current HIP device)."""

    def __init__(self, case, handle=None):
        self.lib = load()
        self.case = case
        self.owned = handle is None
        if handle is not None:          # a case handle owned by a Batch
            self.h = handle
            return
        self.h = C.c_void_p()
        d = case.bind()
        rc = self.lib.hugks_create(C.byref(d), C.byref(self.h))
        if rc != OK:
            raise HugksError(rc, self.lib.hugks_last_error(None).decode())

    def _check(self, rc):
        if rc != OK:
            raise HugksError(rc, self.lib.hugks_last_error(self.h).decode())

    def close(self):
        if self.h and self.owned:
            self.lib.hugks_destroy(self.h)
        self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def step(self):
        info = StepInfo()
        self._check(self.lib.hugks_step(self.h, C.byref(info)))
        return info

    def advance(self, max_steps):
        info = StepInfo()
        self._check(self.lib.hugks_advance(self.h, int(max_steps), C.byref(info)))
        return info

    def advance_until(self, max_steps, t_stop):
        info = StepInfo()
        self._check(self.lib.hugks_advance_until(self.h, int(max_steps), float(t_stop),
                                                 C.byref(info)))
        return info

    def macro(self, with_q=True):
        d = self.case.desc
        w = np.empty((self.case.ncell, d.dim + 2))
        q = np.empty((self.case.ncell, d.dim)) if with_q else None
        self._check(self.lib.hugks_get_macro(self.h, _ptr(w), _ptr(q)))
        return (w, q) if with_q else w

    def macro_noalloc(self, with_q=True):
        """As macro(), through hugks_get_macro_noalloc (no device allocation;
        q staged in the library's static device array)."""
        d = self.case.desc
        w = np.empty((self.case.ncell, d.dim + 2))
        q = np.empty((self.case.ncell, d.dim)) if with_q else None
        self._check(self.lib.hugks_get_macro_noalloc(self.h, _ptr(w), _ptr(q)))
        return (w, q) if with_q else w

    def write_blocks_async(self, path, header, per_line, blocks):
        """Queues a macro*.dat-style file on the handle's writer thread
        (hugks_write_blocks_async): header text, then each block (a 1D array)
        as records of per_line ES23.16 fields."""
        self._check(write_blocks(path, header, per_line, blocks, self.lib, self.h))

    def write_wait(self):
        self._check(self.lib.hugks_write_wait(self.h))

    def get_f(self):
        n = self.case.ncell * self.case.nn
        hf = np.empty(n)
        bf = np.empty(n)
        self._check(self.lib.hugks_get_f(self.h, _ptr(hf), _ptr(bf)))
        return hf.reshape(self.case.ncell, self.case.nn), bf.reshape(self.case.ncell, self.case.nn)

    def set_f(self, hf, bf, w=None):
        hf = np.ascontiguousarray(hf, dtype=np.float64)
        bf = np.ascontiguousarray(bf, dtype=np.float64)
        w = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
        self._check(self.lib.hugks_set_f(self.h, _ptr(hf), _ptr(bf), _ptr(w)))

    def debug_fetch(self, which):
        """Intermediate array of the last step (ARR_* ids), dense FP64."""
        d = self.case.desc
        two = d.dim == 2
        nvf = (d.nx + 1) * d.ny if two else d.nx + 1
        nhf = d.nx * (d.ny + 1) if two else 0
        nn = self.case.nn
        rows = {ARR_SHX: (self.case.ncell, nn), ARR_SHY: (self.case.ncell, nn),
                ARR_SBX: (self.case.ncell, nn), ARR_SBY: (self.case.ncell, nn),
                ARR_VFH: (nvf, nn), ARR_VFB: (nvf, nn), ARR_HFH: (nhf, nn), ARR_HFB: (nhf, nn),
                ARR_VFLUX: (nvf, 4), ARR_HFLUX: (nhf, 4)}[which]
        out = np.empty(rows)
        self._check(self.lib.hugks_debug_fetch(self.h, int(which), _ptr(out)))
        return out

    def checkpoint(self, path):
        """Writes the whole solver state to path (hugks_checkpoint_write)."""
        self._check(self.lib.hugks_checkpoint_write(self.h, str(path).encode()))

    def restore(self, path):
        """Loads a state written by checkpoint() into this handle."""
        self._check(self.lib.hugks_checkpoint_read(self.h, str(path).encode()))

    def time(self):
        t = C.c_double()
        it = C.c_int32()
        self._check(self.lib.hugks_get_time(self.h, C.byref(t), C.byref(it)))
        return t.value, it.value

    def set_timing(self, on):
        self._check(self.lib.hugks_set_timing(self.h, 1 if on else 0))

    def reset_kernel_times(self):
        self._check(self.lib.hugks_reset_kernel_times(self.h))

    def kernel_times(self):
        cap = 32
        names = (C.c_char_p * cap)()
        ms = (C.c_double * cap)()
        calls = (C.c_int32 * cap)()
        n = C.c_int32()
        self._check(self.lib.hugks_kernel_times(self.h, cap, names, ms, calls, C.byref(n)))
        return {names[i].decode(): (ms[i], calls[i]) for i in range(min(n.value, cap))}


    def set_schedule(self, flux_kernel, update_kernel):
        self._check(self.lib.hugks_set_schedule(self.h, int(flux_kernel), int(update_kernel)))

    def schedule(self):
        f, u = C.c_int32(), C.c_int32()
        self._check(self.lib.hugks_get_schedule(self.h, C.byref(f), C.byref(u)))
        return f.value, u.value

    def set_update_waves(self, waves):
        self._check(self.lib.hugks_set_update_waves(self.h, int(waves)))

    def update_waves(self):
        w = C.c_int32()
        self._check(self.lib.hugks_get_update_waves(self.h, C.byref(w)))
        return w.value


def tune(case):
    """(flux_kernel, update_kernel, ms_per_step) the autotuner picks for
    the case's shape."""
    lib = load()
    f, u, ms = C.c_int32(), C.c_int32(), C.c_double()
    d = case.bind()
    rc = lib.hugks_tune(C.byref(d), C.byref(f), C.byref(u), C.byref(ms))
    if rc != OK:
        raise HugksError(rc, lib.hugks_last_error(None).decode())
    return f.value, u.value, ms.value


def tune_tiles(case):
    """(flux_kernel, update_kernel, update_waves, ms_per_step) the
    autotuner picks for the case's shape, tile included."""
    lib = load()
    f, u, w, ms = C.c_int32(), C.c_int32(), C.c_int32(), C.c_double()
    d = case.bind()
    rc = lib.hugks_tune_tiles(C.byref(d), C.byref(f), C.byref(u), C.byref(w), C.byref(ms))
    if rc != OK:
        raise HugksError(rc, lib.hugks_last_error(None).decode())
    return f.value, u.value, w.value, ms.value


class Batch:
    """Independent cases advanced together by one graph replay per step."""

    def __init__(self, case_list, autotune=False):
        self.lib = load()
        self.cases = list(case_list)
        descs = (Desc * len(self.cases))()
        for i, c in enumerate(self.cases):
            descs[i] = c.bind()
        self._descs = descs
        self.b = C.c_void_p()
        rc = self.lib.hugks_batch_create(len(self.cases), descs,
                                         BATCH_AUTOTUNE if autotune else 0, C.byref(self.b))
        if rc != OK:
            raise HugksError(rc, self.lib.hugks_batch_last_error(None).decode())
        self.solvers = []
        for i, c in enumerate(self.cases):
            h = C.c_void_p()
            self._check(self.lib.hugks_batch_handle(self.b, i, C.byref(h)))
            self.solvers.append(Solver(c, handle=h))

    def _check(self, rc):
        if rc != OK:
            raise HugksError(rc, self.lib.hugks_batch_last_error(self.b).decode())

    @property
    def grouped(self):
        """True when grouped launches (one per kernel type) run every case."""
        return self.grouped_cases == len(self.solvers)

    @property
    def grouped_cases(self):
        """Number of cases stepped by the grouped launches."""
        return int(self.lib.hugks_batch_grouped(self.b))

    def advance(self, max_steps):
        infos = (StepInfo * len(self.cases))()
        self._check(self.lib.hugks_batch_advance(self.b, int(max_steps), infos))
        return list(infos)

    def set_timing(self, on):
        """Eager steps with the grouped launches timed (a measurement aid)."""
        self._check(self.lib.hugks_batch_set_timing(self.b, 1 if on else 0))

    def reset_kernel_times(self):
        self._check(self.lib.hugks_batch_reset_kernel_times(self.b))

    def kernel_times(self):
        cap = 32
        names = (C.c_char_p * cap)()
        ms = (C.c_double * cap)()
        calls = (C.c_int32 * cap)()
        n = C.c_int32()
        self._check(self.lib.hugks_batch_kernel_times(self.b, cap, names, ms, calls,
                                                      C.byref(n)))
        return {names[i].decode(): (ms[i], calls[i]) for i in range(min(n.value, cap))}

    def close(self):
        if self.b:
            for s in self.solvers:
                s.h = C.c_void_p()
            self.lib.hugks_batch_destroy(self.b)
            self.b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def device_info():
    return load().hugks_device_info().decode()
