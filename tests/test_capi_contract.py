"""The C-ABI boundary without a GPU: libhugks.so loads, exports every
function include/hugks.h declares, the ctypes mirrors have the C layout,
and the error paths that need no device return the header's codes."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from hugks import capi, cases, namelist

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hugks.h")


def header_text():
    with open(HEADER) as f:
        return f.read()


def declared_functions():
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w[\w\s\*]*?\b(hugks_\w+)\s*\(",
                                 header_text(), flags=re.M)))


def test_header_declares_the_bound_functions():
    assert declared_functions() == sorted(capi.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIBPATH], check=True,
                         capture_output=True, text=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    for name in declared_functions():
        assert name in exported, name
        assert hasattr(lib, name)


def test_header_constants_match_binding():
    text = header_text()
    consts = dict((k, int(v)) for k, v in re.findall(r"#define\s+(HUGKS_\w+)\s+(-?\d+)\b", text))
    assert consts["HUGKS_ABI_VERSION"] == capi.ABI_VERSION
    for name in ("OK", "E_ARG", "E_NOMEM", "E_DEVICE", "E_UNSUPPORTED", "E_NUMERIC", "E_WALL",
                 "E_NOHIP", "FP64", "MIXED_FP32", "MIXED_BF16X3", "BC_PERIODIC", "BC_SPECULAR",
                 "BC_DIFFUSE", "TAU_MU_OVER_RHO", "TAU_VHS", "DT_2D_GENERAL", "DT_1D",
                 "FORCE_HERMITE_SYMMETRIC", "FORCE_HERMITE_SYMMETRIC_B", "FORCE_HERMITE_LEGACY",
                 "STOP_NONE", "STOP_MAXTIME", "STOP_RESIDUAL", "STOP_TIME", "STOP_STEADY", "STOP_ERROR",
                 "ARR_SHX", "ARR_HFLUX"):
        assert consts["HUGKS_" + name] == getattr(capi, name), name
    assert consts["HUGKS_DT_2D_WALL_NORMAL"] == capi.DT_2D_WALL_NORMAL


LAYOUT_C = r"""
#include <stddef.h>
#include <stdio.h>
#include "hugks.h"
#define F(s, m) printf(#s " " #m " %zu\n", offsetof(s, m))
int main(void) {
    printf("hugks_desc sizeof %zu\n", sizeof(hugks_desc));
    printf("hugks_step_info sizeof %zu\n", sizeof(hugks_step_info));
    %FIELDS%
    return 0;
}
"""


def test_struct_layout_matches_c(tmp_path):
    lines = []
    for sname, cls in (("hugks_desc", capi.Desc), ("hugks_step_info", capi.StepInfo)):
        for fname, _ in cls._fields_:
            lines.append(f"F({sname}, {fname});")
    src = tmp_path / "layout.c"
    src.write_text(LAYOUT_C.replace("%FIELDS%", "\n    ".join(lines)))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        s, m, v = line.split()
        got[(s, m)] = int(v)
    assert got[("hugks_desc", "sizeof")] == C.sizeof(capi.Desc)
    assert got[("hugks_step_info", "sizeof")] == C.sizeof(capi.StepInfo)
    for sname, cls in (("hugks_desc", capi.Desc), ("hugks_step_info", capi.StepInfo)):
        for fname, _ in cls._fields_:
            assert got[(sname, fname)] == getattr(cls, fname).offset, (sname, fname)


def small_case():
    p = dict(namelist.DEFAULTS_2D)
    p.update(nx=4, ny=8, nv=12, Nm=6)
    return cases.build_2d(p)


def create(desc):
    lib = capi.load()
    h = C.c_void_p()
    rc = lib.hugks_create(C.byref(desc), C.byref(h))
    msg = lib.hugks_last_error(None).decode()
    if h:
        lib.hugks_destroy(h)
    return rc, msg


BAD = [
    ("abi_version", 99), ("dim", 3), ("nx", 0), ("ny", -1), ("nv", 1), ("nm", -1),
    ("precision", 7), ("tau_model", 5), ("dt_rule", 9), ("force_variant", 4), ("dx", 0.0),
    ("dy", -1.0), ("cfl", 0.0), ("prandtl", 0.0), ("smv", 0.0),
]


@pytest.mark.parametrize("field,value", BAD, ids=[b[0] for b in BAD])
def test_bad_descriptor_is_rejected_before_any_device_work(field, value):
    case = small_case()
    d = case.bind()
    setattr(d, field, value)
    rc, msg = create(d)
    assert rc == capi.E_ARG, (rc, msg)
    assert msg


def test_bad_boundary_code_is_rejected():
    case = small_case()
    d = case.bind()
    d.bc[2] = 7
    assert create(d)[0] == capi.E_ARG


def test_missing_arrays_are_rejected():
    case = small_case()
    d = case.bind()
    d.weight = None
    assert create(d)[0] == capi.E_ARG


def test_null_arguments():
    lib = capi.load()
    assert lib.hugks_create(None, None) == capi.E_ARG
    assert lib.hugks_step(None, None) == capi.E_ARG
    assert lib.hugks_destroy(None) in (capi.OK, capi.E_ARG)
    assert lib.hugks_last_error(None) is not None


def test_valid_descriptor_without_device(have_no_gpu):
    """A valid descriptor on a host without a HIP device is refused with
    E_NOHIP: the library has no CPU path to fall back on."""
    case = small_case()
    rc, msg = create(case.bind())
    assert rc == capi.E_NOHIP, (rc, msg)


@pytest.fixture
def have_no_gpu():
    info = capi.device_info()
    if not info.startswith("no HIP device"):
        pytest.skip("a HIP device is present")
    return info


def test_bf16x3_refused_in_1d():
    """HUGKS_MIXED_BF16X3 names the split-product 2D Hermite contraction; a
    1D case has none, and the library refuses the mode instead of running
    another one (validation before any device call, so also without a GPU)."""
    p = dict(namelist.DEFAULTS_1D)
    p.update(nx=16, nv=12)
    case = cases.build_1d(p)
    d = case.bind()
    d.precision = capi.MIXED_BF16X3
    rc, msg = create(d)
    assert rc == capi.E_UNSUPPORTED, (rc, msg)
    assert "BF16X3" in msg and "2D" in msg


def header_defines():
    return {k: int(v) for k, v in re.findall(r"^#define\s+(HUGKS_\w+)\s+(-?\d+)\b", header_text(),
                                             flags=re.M)}


def test_fortran_module_constants_match_the_header():
    """The Fortran binding's named constants carry the header's values, and
    it names every schedule the header defines."""
    with open(os.path.join(ROOT, "hermite-ugks_amd", "driver", "hugks_capi.f90")) as f:
        src = f.read()
    fort = {k: int(v) for k, v in re.findall(r"\b(HUGKS_\w+)\s*=\s*(-?\d+)\b", src)}
    hdr = header_defines()
    common = sorted(set(fort) & set(hdr))
    assert common
    for k in common:
        assert fort[k] == hdr[k], k
    for k in hdr:
        if k.startswith(("HUGKS_FLUX_", "HUGKS_UPDATE_")):
            assert k in fort, k


def test_python_schedule_constants_match_the_header():
    hdr = header_defines()
    for k, v in hdr.items():
        if k.startswith(("HUGKS_FLUX_", "HUGKS_UPDATE_")):
            assert getattr(capi, k[len("HUGKS_"):]) == v, k
