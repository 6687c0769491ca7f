"""The macro*.dat writer of the library (hugks_write_blocks_async with no
handle: the file is written before the call returns), on the CPU.

- Every macro*.dat file the reference's own main programs wrote
  (tests/golden/macro/*.npz, from oracle/_ref through
  tests/golden/make_golden.py), 2D and 1D: its header and its numbers,
  parsed back to doubles (17 significant digits identify a double) and
  handed to the writer as the blocks of the file, come out byte for byte
  as the reference wrote them (2d/src/output2d.f90:38-55,
  1d/src/output.f90:89-99).
- The edge cases of the Fortran ES23.16 edit descriptor, written by the
  compiler of the reference build (tests/golden/es/es23_16_values.txt,
  from make_es_fixture.sh): signed zero, three-digit exponents without
  the E, the smallest subnormal, an exact rounding tie (2**-25: ties to
  even), NaN and infinities, and an empty array (one empty record).
- Argument errors and an unwritable path.
"""
import glob
import os
import re

import numpy as np
import pytest

from hugks import capi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MACRO = sorted(glob.glob(os.path.join(GOLDEN, "macro", "*.npz")))
NUM = re.compile(r"[-+]?\d\.\d{16}(?:E[-+]\d{2}|[-+]\d{3})|[-+]?(?:NaN|Inf)")


def parse_es(tok):
    """A Fortran ES23.16 field back to its double."""
    if tok.endswith("NaN"):
        return float("nan")
    if tok.endswith("Inf"):
        return float("-inf") if tok.startswith("-") else float("inf")
    if "E" not in tok:                       # three-digit exponent, E dropped
        m = re.match(r"([-+]?\d\.\d{16})([-+]\d{3})$", tok)
        return float(m.group(1) + "E" + m.group(2))
    return float(tok)


def _load():
    capi.load()


def macro_files():
    out = []
    for p in MACRO:
        z = np.load(p)
        for name in z.files:
            out.append((os.path.basename(p)[:-4] + "/" + name, str(z[name])))
    return out


MACRO_FILES = macro_files()


@pytest.mark.parametrize("name,text", MACRO_FILES, ids=[n for n, _ in MACRO_FILES])
def test_reference_macro_files_rewritten_byte_for_byte(tmp_path, name, text):
    _load()
    lines = text.split("\n")
    header = lines[0] + "\n" + lines[1] + "\n"
    zone = lines[1]
    if " J =" in zone:                       # 2D: node coordinates, then 11 cell fields
        ni, nj = (int(x) for x in re.findall(r"[IJ] =\s*(\d+)", zone))
        per_line, lens = 11, [ni * nj] * 2 + [(ni - 1) * (nj - 1)] * 11
    else:                                    # 1D: x, then 9 cell fields
        n = int(re.findall(r"I =\s*(\d+)", zone)[0])
        per_line, lens = 9, [n] * 10
    vals = [parse_es(t) for t in NUM.findall("\n".join(lines[2:]))]
    assert len(vals) == sum(lens), (len(vals), sum(lens))
    blocks, o = [], 0
    for n in lens:
        blocks.append(np.array(vals[o:o + n]))
        o += n
    out = tmp_path / "macro.dat"
    rc = capi.write_blocks(out, header, per_line, blocks)
    assert rc == capi.OK, capi.load().hugks_last_error(None)
    got = out.read_bytes()
    assert got == text.encode(), name


def test_es_edge_cases_as_the_fortran_runtime_writes_them(tmp_path):
    _load()
    ref = open(os.path.join(GOLDEN, "es", "es23_16_values.txt")).read()
    recs = ref.split("\n")
    toks = NUM.findall("\n".join(recs[:5]))
    v = np.array([parse_es(t) for t in toks])
    assert v.size == 40
    assert np.signbit(v[1]) and v[18] == 5e-324 and v[19] == 2.0 ** -25
    out = tmp_path / "es.txt"
    rc = capi.write_blocks(out, "", 9, [v, v[:0], v[:9]])
    assert rc == capi.OK
    assert out.read_text() == ref


def test_writer_argument_errors(tmp_path):
    lib = capi.load()
    assert capi.write_blocks(tmp_path / "a.dat", "h\n", 0, [np.ones(3)]) == capi.E_ARG
    assert capi.write_blocks(tmp_path / "no" / "such" / "dir.dat", "h\n", 9,
                             [np.ones(3)]) == capi.E_IO
    assert "dir.dat" in lib.hugks_last_error(None).decode()
    assert lib.hugks_write_wait(None) == capi.E_ARG
