"""macro*.dat output through a solver handle's writer thread
(hugks_write_blocks_async / hugks_write_wait), the path the Fortran
drivers take (the files themselves are checked against the reference's in
test_gpu_driver.py, the formatting byte for byte in test_block_writer.py).

- Files queued while the solver keeps stepping are written in call order,
  each byte for byte what the synchronous writer (no handle) makes of the
  same blocks, which were copied at the call (the caller's buffer is
  overwritten right after queueing).
- An unwritable path is reported by the next wait as HUGKS_E_IO naming the
  file, and only once.
- Files still queued when the handle is destroyed are written.
"""
import numpy as np
import pytest

from hugks import capi, cases, namelist

pytestmark = pytest.mark.gpu


def small_case():
    p = dict(namelist.DEFAULTS_2D)
    p.update(nx=6, ny=8, nv=12, Nm=6)
    return cases.build_2d(p, capi.FP64)


def blocks_of(w, q):
    return [w[:, 0].copy(), w[:, 1].copy(), w[:, 2].copy(), w[:, 3].copy(), q[:, 0].copy()]


def test_queued_files_equal_synchronous_writes(have_gpu, tmp_path):
    header = " VARIABLES = A, B\nZONE I =    6, J =    8\n"
    expected = []
    with capi.Solver(small_case()) as s:
        for k in range(3):
            s.advance(2)
            w, q = s.macro()
            buf = np.concatenate(blocks_of(w, q))
            s.write_blocks_async(tmp_path / f"macro{k}.dat", header, 11,
                                 [buf[i * 48:(i + 1) * 48] for i in range(5)])
            buf[:] = -1.0        # the writer holds its own copy
            ref = tmp_path / f"ref{k}.dat"
            assert capi.write_blocks(ref, header, 11, blocks_of(w, q)) == capi.OK
            expected.append(ref.read_bytes())
        s.advance(3)             # steps go on while the files are written
        s.write_wait()
    for k in range(3):
        assert (tmp_path / f"macro{k}.dat").read_bytes() == expected[k], k


def test_write_error_reported_by_wait(have_gpu, tmp_path):
    with capi.Solver(small_case()) as s:
        s.write_blocks_async(tmp_path / "missing" / "macro1.dat", "h\n", 9, [np.ones(4)])
        with pytest.raises(capi.HugksError) as e:
            s.write_wait()
        assert e.value.code == capi.E_IO and "macro1.dat" in str(e.value)
        s.write_wait()           # reported once
        s.write_blocks_async(tmp_path / "macro2.dat", "h\n", 9, [np.ones(4)])
    assert (tmp_path / "macro2.dat").read_text().startswith("h\n 1.0000000000000000E+00")
