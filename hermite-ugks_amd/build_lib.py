"""Builds libhugks.so (gfx950) in-tree with hipcc.

Usage: python hermite-ugks_amd/build_lib.py [--jobs N]

Every translation unit under csrc/ is compiled separately (in parallel)
and linked into hermite-ugks_amd/libhugks.so.  Packed FP32 instructions
are disabled for device code.  The build fails when the compiler's
resource-usage remarks report scratch (private) memory, a spilled VGPR or
all 256 VGPRs in use for any kernel.  The Fortran drivers under driver/
(ugks2d_gpu, ugks1d_gpu: the reference's main programs over the C-ABI)
are then built with amdflang and linked against the library.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhugks.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")
DRIVER = os.path.join(HERE, "driver")
ARCH = os.environ.get("HUGKS_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def compile_one(src):
    obj = os.path.join(OBJ, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC, *FLAGS, "-Rpass-analysis=kernel-resource-usage", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r.returncode, r.stderr


def scratch_kernels(stderr):
    """(kernel, what) of every kernel whose resource-usage remarks report a
    private segment (register spills or stack), spilled VGPRs, or the
    whole 256-VGPR file in use (where the compiler starts keeping spilled
    values in AGPRs: the register picture of a flux kernel that once gave
    wrong values and a memory fault).  SGPRs spilled to VGPR lanes are
    allowed: they stay in registers."""
    out, name = [], None
    for ln in stderr.splitlines():
        if "Function Name:" in ln:
            name = ln.split("Function Name:")[1].split()[0]
            continue
        if not name:
            continue
        for key, limit, what in (("ScratchSize [bytes/lane]:", 0, "bytes/lane of scratch"),
                                 ("VGPRs Spill:", 0, "VGPRs spilled"),
                                 (" VGPRs:", 255, "VGPRs (the whole file)")):
            if key in ln:
                n = int(ln.split(key)[1].split()[0])
                if n > limit:
                    out.append((name, f"{n} {what}"))
    return out


def build(jobs=None, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), os.cpu_count() or 4, 8)
    objs = []
    errors = []
    spills = []
    with cf.ThreadPoolExecutor(jobs) as ex:
        for src, obj, rc, err in ex.map(compile_one, srcs):
            # keep only real diagnostics (the host pass does not know the
            # device-only target feature and says so; the resource-usage
            # remarks are read below)
            lines = [ln for ln in err.splitlines()
                     if "packed-fp32-ops" not in ln and "kernel-resource-usage" not in ln
                     and ln.strip()]
            if rc != 0:
                errors.append(f"{src}:\n" + "\n".join(lines))
            elif verbose and lines:
                print("\n".join(lines), file=sys.stderr)
            spills += [(os.path.basename(src), k, n) for k, n in scratch_kernels(err)]
            objs.append(obj)
    if errors:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errors))
    # no shipped kernel may use scratch (private) memory or spill vector
    # registers: a spill is a silent slowdown and a class of fault this
    # library rules out
    if spills:
        raise RuntimeError("kernels that spill or use scratch memory:\n" +
                           "\n".join(f"  {f}: {k} {n}" for f, k, n in spills))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, LIB)
    build_drivers()
    return LIB


def build_drivers():
    """ugks2d_gpu and ugks1d_gpu from driver/*.f90 (ISO_C_BINDING module
    first), linked to libhugks.so with an rpath to this directory."""
    if not os.path.exists(FC):
        return []
    mod = os.path.join(OBJ, "fmod")
    os.makedirs(mod, exist_ok=True)
    common = ["hugks_capi.f90", "hugks_quadrature.f90"]
    objs = []
    for f in common:
        o = os.path.join(mod, f[:-4] + ".o")
        r = subprocess.run([FC, "-O2", "-J", mod, "-c", os.path.join(DRIVER, f), "-o", o],
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang {f} failed:\n{r.stderr}")
        objs.append(o)
    out = []
    for prog in ("ugks2d_gpu", "ugks1d_gpu"):
        exe = os.path.join(DRIVER, prog)
        r = subprocess.run([FC, "-O2", "-J", mod, os.path.join(DRIVER, prog + ".f90"), *objs,
                            "-L", HERE, "-lhugks", "-Wl,-rpath," + HERE, "-o", exe],
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang {prog} failed:\n{r.stderr}")
        out.append(exe)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.v))
