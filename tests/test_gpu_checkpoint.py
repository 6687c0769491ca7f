"""Checkpoint / restore (hugks_checkpoint_write / _read): a handle restored
from a checkpoint continues bitwise as the handle that wrote it (W, f and
the loop state), in both precisions; a checkpoint of another shape,
Hermite order or precision, or of a run with other weights, forces or
physical constants, is refused with HUGKS_E_ARG, while loop controls
(maxtime, cfl) may change."""
import numpy as np
import pytest

from hugks import capi, cases
from test_gpu_parity import rt, poiseuille

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [capi.FP64, capi.MIXED_FP32])
@pytest.mark.parametrize("name", ["rt", "poiseuille"])
def test_restore_continues_bitwise(have_gpu, tmp_path, prec, name):
    p = rt(nx=6, ny=8, nv=14, nm=8) if name == "rt" else poiseuille(nx=4, ny=10, nv=12)
    case = cases.build_2d(p, prec, rt_variant="cartesian" if name == "rt" else "reference", seed=2)
    path = tmp_path / "state.ckpt"
    print(name, "prec", prec, flush=True)
    with capi.Solver(case) as a:
        a.advance(5)
        a.checkpoint(path)
        ia = a.advance(4)
        wa = a.macro(with_q=False)
        ha, ba = a.get_f()
        ta = a.time()
    with capi.Solver(case) as b:
        b.restore(path)
        ib = b.advance(4)
        wb = b.macro(with_q=False)
        hb, bb = b.get_f()
        tb = b.time()
    assert np.array_equal(wa, wb) and np.array_equal(ha, hb) and np.array_equal(ba, bb)
    assert ta == tb and ia.iter == ib.iter and ia.simtime == ib.simtime and ia.dt == ib.dt
    assert list(ia.res) == list(ib.res)


def test_restore_refuses_other_shape(have_gpu, tmp_path):
    path = tmp_path / "state.ckpt"
    with capi.Solver(cases.build_2d(poiseuille(nx=4, ny=10, nv=12), capi.FP64)) as a:
        a.advance(1)
        a.checkpoint(path)
    for other in (cases.build_2d(poiseuille(nx=4, ny=12, nv=12), capi.FP64),
                  cases.build_2d(poiseuille(nx=4, ny=10, nv=12), capi.MIXED_FP32),
                  cases.build_2d(poiseuille(nx=4, ny=10, nv=12, nm=5), capi.FP64)):
        with capi.Solver(other) as b:
            with pytest.raises(capi.HugksError, match="differs"):
                b.restore(path)
    # same shape, other run: force, Knudsen number (collision time), wall
    for other in (poiseuille(nx=4, ny=10, nv=12, phi=0.07),
                  poiseuille(nx=4, ny=10, nv=12, kn=0.2),
                  poiseuille(nx=4, ny=10, nv=12, tw=(1.0, 1.1))):
        with capi.Solver(cases.build_2d(other, capi.FP64)) as b:
            with pytest.raises(capi.HugksError, match="written by a run with other"):
                b.restore(path)
    # loop controls may differ
    longer = dict(poiseuille(nx=4, ny=10, nv=12), maxtime=500.0, cfl=0.6)
    with capi.Solver(cases.build_2d(longer, capi.FP64)) as b:
        b.restore(path)
    with capi.Solver(cases.build_2d(poiseuille(nx=4, ny=10, nv=12), capi.FP64)) as b:
        (tmp_path / "junk").write_bytes(b"not a checkpoint")
        with pytest.raises(capi.HugksError, match="not a checkpoint"):
            b.restore(tmp_path / "junk")
