"""Reader for the reference's Fortran ``&params`` namelist files
(2d/src/config.f90:47-48, 1d/src/config.f90:48-50) and the optional
``&gpu`` group that follows it in this project's input files.

Defaults are the reference's CONFIG module defaults, so a file that sets
only some names behaves as it does for the reference.
"""
import re

DEFAULTS_2D = dict(  # 2d/src/config.f90:12-40
    nx=3, ny=100, xlength=0.1, ylength=1.0, nv=16, vmax=5.0, Tref=1.0,
    quad_type="gauss_hermite", force_type="hermite_symmetric", Nm=7, cfl=0.8,
    maxtime=200.0, case_id=5, knudsen=2e-2, phi=0.1, TwL=1.0, TwR=1.0,
)
DEFAULTS_1D = dict(  # 1d/src/config.f90:11-45
    nx=600, xlength=1.0, nv=100, vmax=5.0, Tref=1.0, quad_type="newton_cotes",
    force_type="hermite", Nm=6, cfl=0.5, maxtime=0.2, steady_tol=1e-9, mass_tol=1e-10,
    steady_steps=1000, case_id=3, knudsen=0.01, phi=-1.0, TwL=1.0, TwR=1.1, refpfile="",
)
DEFAULTS_GPU = dict(precision="fp64", max_steps=0, rt_variant="reference", seed=0)


def _value(tok):
    t = tok.strip()
    if not t:
        return None
    if t[0] in "'\"":
        return t[1:-1] if len(t) >= 2 and t[-1] == t[0] else t[1:]
    low = t.lower()
    if low in (".true.", "t", ".t."):
        return True
    if low in (".false.", "f", ".f."):
        return False
    try:
        if re.fullmatch(r"[+-]?\d+", t):
            return int(t)
        return float(re.sub(r"[dD]", "e", t))
    except ValueError:
        return t


def parse_groups(text):
    """Returns {group_name: {name: value}} for every &group ... / in text."""
    # strip comments outside quotes
    lines = []
    for ln in text.splitlines():
        out, q = [], None
        for ch in ln:
            if q:
                out.append(ch)
                if ch == q:
                    q = None
            elif ch in "'\"":
                q = ch
                out.append(ch)
            elif ch == "!":
                break
            else:
                out.append(ch)
        lines.append("".join(out))
    body = "\n".join(lines)
    groups = {}
    for m in re.finditer(r"&(\w+)(.*?)(?<!['\"])/", body, re.S):
        name = m.group(1).lower()
        items = {}
        content = m.group(2)
        for km in re.finditer(r"(\w+)\s*=\s*('[^']*'|\"[^\"]*\"|[^,\n=]+?)\s*(?=,|\n|$|\w+\s*=)",
                              content):
            items[km.group(1)] = _value(km.group(2))
        groups[name] = items
    return groups


def _canon(d, defaults):
    out = dict(defaults)
    lower = {k.lower(): k for k in defaults}
    for k, v in d.items():
        key = lower.get(k.lower(), k)
        out[key] = v
    return out


def read(path, dim):
    """Reads &params (with the reference defaults for dim) and &gpu."""
    with open(path) as f:
        groups = parse_groups(f.read())
    params = _canon(groups.get("params", {}), DEFAULTS_2D if dim == 2 else DEFAULTS_1D)
    gpu = _canon(groups.get("gpu", {}), DEFAULTS_GPU)
    return params, gpu
