"""The MFMA node update (HUGKS_UPDATE_MFMA_NODE: force and collision terms
on matrix tiles, each node updated from its tile accumulator) against the
FP64 oracle: FP64 within 1e-10, mixed within 1e-5, for every compiled
wavefronts-per-cell instance at the compiled-in grid sizes and at grid
sizes that take the run-time instances (odd, and not multiples of 4), with
Hermite orders below 3 (the heat flux then comes from coefficients that the
force does not use), up to the largest supported (14), with unequal walls
and with weights that are not a tensor product.

The same for the node update that also forms the interior faces' node
fluxes (HUGKS_UPDATE_MFMA_FACE, with the cell-sum flux): every instance,
unequal diffuse walls, specular walls, a grid periodic in both directions
(every face interior, the neighbours wrapping around), and with the split
flux, weights that are not a tensor product."""
import numpy as np
import pytest

from hugks import capi, cases
import oracle_py
from metrics import field_errors
from test_gpu_parity import poiseuille, rt

pytestmark = pytest.mark.gpu

TOL = {capi.FP64: 1e-10, capi.MIXED_FP32: 1e-5}
# (nv, waves per cell) with an instance: compiled-in 16..40, run-time others
GRID_WAVES = [(16, 1), (20, 4), (20, 2), (24, 4), (28, 2), (32, 4), (32, 2), (40, 3), (40, 5),
              (40, 9), (13, 1), (22, 4), (30, 2), (36, 3), (45, 5)]


NODE = (capi.FLUX_SPLIT, capi.UPDATE_MFMA_NODE)
FACE = (capi.FLUX_CELLSUM, capi.UPDATE_MFMA_FACE)
FACE_SPLIT = (capi.FLUX_SPLIT, capi.UPDATE_MFMA_FACE)


def check(case, steps, waves=0, sched=NODE):
    res = []
    o = oracle_py.Oracle(case.with_precision(capi.FP64))
    o.step(steps)
    wr = o.get()[0]
    for prec in (capi.FP64, capi.MIXED_FP32):
        c = case.with_precision(prec)
        with capi.Solver(c) as s:
            s.set_schedule(*sched)
            if waves:
                s.set_update_waves(waves)
            assert s.schedule() == sched
            s.advance(steps)
            w = s.macro(with_q=False)
        err = field_errors(w, wr, thermal=True)
        print(f"  prec={prec} waves={waves or 'default'} W err {err}", flush=True)
        assert np.all(err <= TOL[prec]), (prec, err)
        res.append(err)
    return res


@pytest.mark.parametrize("nv,waves", GRID_WAVES, ids=[f"nv{a}-w{b}" for a, b in GRID_WAVES])
def test_node_update_instances(have_gpu, nv, waves):
    nm = min(10, 2 * nv - 2)
    print(f"RT 6x8 nv={nv} Nm={nm} waves={waves}", flush=True)
    check(cases.build_2d(rt(nx=6, ny=8, nv=nv, nm=nm), capi.FP64, rt_variant="cartesian",
                         seed=nv), 5, waves)


@pytest.mark.parametrize("nm", [0, 1, 2, 3, 14])
def test_node_update_hermite_orders(have_gpu, nm):
    print(f"RT 6x8 nv=24 Nm={nm}", flush=True)
    check(cases.build_2d(rt(nx=6, ny=8, nv=24, nm=nm), capi.FP64, rt_variant="cartesian",
                         seed=4), 5)


def test_node_update_channel_unequal_walls(have_gpu):
    print("channel 4x10 nv=20 TwR=1.3 legacy force", flush=True)
    check(cases.build_2d(poiseuille(nx=4, ny=10, nv=20, nm=8, force="hermite", tw=(1.0, 1.3)),
                         capi.FP64), 8)


def test_node_update_weights_not_a_tensor_product(have_gpu):
    case = cases.build_2d(rt(nx=6, ny=8, nv=16, nm=8), capi.FP64, rt_variant="cartesian", seed=6)
    w = case.arrays["weight"]
    w *= 1.0 + 1e-3 * np.random.default_rng(3).uniform(-1.0, 1.0, w.shape)
    print("perturbed weights nv=16", flush=True)
    check(case, 5)


def test_set_update_waves_refuses_missing_instances(have_gpu):
    case = cases.build_2d(rt(nx=4, ny=4, nv=40, nm=8), capi.MIXED_FP32, rt_variant="cartesian")
    with capi.Solver(case) as s:
        assert s.update_waves() == 3
        for bad in (1, 2, 4):
            with pytest.raises(capi.HugksError) as ei:
                s.set_update_waves(bad)
            assert ei.value.code == capi.E_UNSUPPORTED
        s.set_update_waves(9)
        assert s.update_waves() == 9
        s.set_update_waves(0)
        assert s.update_waves() == 3


@pytest.mark.parametrize("nv,waves", GRID_WAVES, ids=[f"nv{a}-w{b}" for a, b in GRID_WAVES])
def test_face_update_instances(have_gpu, nv, waves):
    nm = min(10, 2 * nv - 2)
    print(f"face-forming update: RT 6x8 nv={nv} Nm={nm} waves={waves}", flush=True)
    check(cases.build_2d(rt(nx=6, ny=8, nv=nv, nm=nm), capi.FP64, rt_variant="cartesian",
                         seed=nv), 5, waves, FACE)


def test_face_update_channel_unequal_walls(have_gpu):
    print("face-forming update: channel 4x10 nv=20 TwR=1.3 legacy force", flush=True)
    check(cases.build_2d(poiseuille(nx=4, ny=10, nv=20, nm=8, force="hermite", tw=(1.0, 1.3)),
                         capi.FP64), 8, sched=FACE)


@pytest.mark.parametrize("sched", [NODE, FACE, FACE_SPLIT], ids=["node", "face", "face-split"])
def test_doubly_periodic(have_gpu, sched):
    case = cases.build_2d(poiseuille(nx=6, ny=8, nv=16, nm=8), capi.FP64)
    case.desc.bc[:] = [capi.BC_PERIODIC] * 4
    print(f"periodic in x and y 6x8 nv=16 sched={sched}", flush=True)
    check(case, 6, sched=sched)


def test_face_update_weights_not_a_tensor_product(have_gpu):
    case = cases.build_2d(rt(nx=6, ny=8, nv=16, nm=8), capi.FP64, rt_variant="cartesian", seed=6)
    w = case.arrays["weight"]
    w *= 1.0 + 1e-3 * np.random.default_rng(3).uniform(-1.0, 1.0, w.shape)
    print("face-forming update, split flux, perturbed weights nv=16", flush=True)
    check(case, 5, sched=FACE_SPLIT)


def test_default_waves_follow_the_update(have_gpu):
    """The default wavefronts per cell at nv = 40: three for both node
    updates in FP32 storage and for the plain node update in FP64, five
    for the face-forming update in FP64 (its faster tile there)."""
    case = cases.build_2d(rt(nx=4, ny=4, nv=40, nm=8), capi.FP64, rt_variant="cartesian")
    with capi.Solver(case) as s:
        assert s.schedule()[1] == capi.UPDATE_MFMA_FACE
        assert s.update_waves() == 5
        s.set_update_waves(0)
        assert s.update_waves() == 5
        s.set_schedule(capi.FLUX_CELLSUM, capi.UPDATE_MFMA_NODE)
        s.set_update_waves(0)
        assert s.update_waves() == 3
    with capi.Solver(case.with_precision(capi.MIXED_FP32)) as s:
        assert s.schedule()[1] == capi.UPDATE_MFMA_FACE
        assert s.update_waves() == 3
