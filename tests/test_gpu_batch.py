"""Batches of independent cases (hugks_batch_*): one graph replay per step
for the whole batch must give, case by case, exactly what the same case
gives when stepped on its own with the same kernel schedule, on every
replay, with ordinary eager work and synchronisations between replays;
the autotuned schedules must still agree with the oracle; each case stops
on its own exit conditions."""
import numpy as np
import pytest

from hugks import capi, cases
import oracle_py
from metrics import field_errors

pytestmark = pytest.mark.gpu


def poiseuille(nx=4, ny=12, nv=12, nm=6, kn=0.1, phi=0.05, tw=(1.0, 1.0), maxtime=200.0):
    return dict(nx=nx, ny=ny, xlength=0.1, ylength=1.0, nv=nv, vmax=5.0, Tref=1.0,
                quad_type="gauss_hermite", force_type="hermite_symmetric", Nm=nm, cfl=0.8,
                maxtime=maxtime, case_id=5, knudsen=kn, phi=phi, TwL=tw[0], TwR=tw[1])


def rt(nx=6, ny=8, nv=12, nm=6):
    return dict(nx=nx, ny=ny, xlength=0.5, ylength=1.0, nv=nv, vmax=5.0, Tref=1.0,
                quad_type="gauss_hermite", force_type="hermite_symmetric", Nm=nm, cfl=0.5,
                maxtime=0.14, case_id=6, knudsen=1e-3, phi=1.0, TwL=1.0, TwR=1.0)


def batch_cases():
    return [
        cases.build_2d(poiseuille(nv=12, ny=10), capi.FP64),
        cases.build_2d(poiseuille(nv=16, ny=14, kn=0.5, phi=0.1, tw=(1.0, 1.2)), capi.MIXED_FP32),
        cases.build_2d(rt(nv=14), capi.FP64, rt_variant="cartesian", seed=4),
        cases.build_2d(poiseuille(nv=20, ny=8, nm=8), capi.MIXED_FP32),
        cases.build_2d(rt(nv=16, nm=8), capi.MIXED_FP32, rt_variant="cartesian", seed=9),
    ]


@pytest.mark.parametrize("autotune,grouped", [(False, True), (False, False), (True, False),
                                              (True, True)])
def test_batch_equals_cases_stepped_alone(have_gpu, monkeypatch, autotune, grouped):
    """FP64 and FP32 cases in one batch: grouped (one plan per storage type)
    or stepped case by case over lane streams, with or without the tuner."""
    cs = batch_cases()
    print(f"batch of {len(cs)} autotune={autotune} grouped={grouped}", flush=True)
    if not grouped:
        monkeypatch.setenv("HUGKS_BATCH_GROUPED", "0")
    with capi.Batch(cs, autotune=autotune) as b:
        assert b.grouped == grouped
        scheds = [s.schedule() for s in b.solvers]
        waves = [s.update_waves() for s in b.solvers]
        print("schedules", scheds, "waves", waves, flush=True)
        singles = []
        for c, sc, nw in zip(cs, scheds, waves):
            s = capi.Solver(c)
            s.set_schedule(*sc)
            s.set_update_waves(nw)
            singles.append(s)
        # 9 and 17 steps: replays of the eight-step graph and a remainder
        for rep, n in enumerate((3, 1, 9, 17)):
            infos = b.advance(n)
            assert [i.steps for i in infos] == [n] * len(cs)
            for k, s in enumerate(singles):
                s.advance(n)
                # eager work on the case's own stream between replays
                wa = s.macro(with_q=False)
                wb = b.solvers[k].macro(with_q=False)
                assert np.array_equal(wa, wb), (rep, k, np.max(np.abs(wa - wb)))
                ha, _ = s.get_f()
                hb, _ = b.solvers[k].get_f()
                assert np.array_equal(ha, hb), (rep, k)
        for s in singles:
            s.close()


def test_batch_autotuned_parity_with_oracle(have_gpu, monkeypatch):
    cs = batch_cases()
    monkeypatch.setenv("HUGKS_BATCH_GROUPED", "0")
    with capi.Batch(cs, autotune=True) as b:
        assert not b.grouped
        b.advance(6)
        ws = [s.macro(with_q=False) for s in b.solvers]
    for c, w in zip(cs, ws):
        o = oracle_py.Oracle(c.with_precision(capi.FP64))
        o.step(6)
        err = field_errors(w, o.get()[0])
        tol = 1e-10 if c.desc.precision == capi.FP64 else 1e-5
        print("case nv", c.desc.nv, "prec", c.desc.precision, "err", err, flush=True)
        assert np.all(err <= tol), err


@pytest.mark.parametrize("grouped", [True, False])
def test_batch_cases_stop_on_their_own(have_gpu, monkeypatch, grouped):
    """Cases with different maxtime: the uncapped advance runs until every
    case has stopped; each stops at its own maxtime after the same number
    of steps as when run alone (grouped launches and per-case graph)."""
    cs = [cases.build_2d(poiseuille(nv=10, ny=8, maxtime=t), capi.FP64) for t in (0.02, 0.05, 0.035)]
    if not grouped:
        monkeypatch.setenv("HUGKS_BATCH_GROUPED", "0")
    with capi.Batch(cs) as b:
        assert b.grouped == grouped
        infos = b.advance(0)
    for c, inf in zip(cs, infos):
        with capi.Solver(c) as s:
            alone = s.advance(0)
        print("maxtime", c.desc.maxtime, "batch steps", inf.steps, "alone", alone.steps, flush=True)
        assert inf.stop_reason == capi.STOP_MAXTIME
        assert inf.steps == alone.steps and inf.iter == alone.iter
        assert inf.simtime == alone.simtime


def grouped_cases(prec):
    return [
        cases.build_2d(poiseuille(nv=12, ny=10), prec),
        cases.build_2d(poiseuille(nv=16, ny=14, kn=0.5, phi=0.1, tw=(1.0, 1.2)), prec),
        cases.build_2d(rt(nv=14), prec, rt_variant="cartesian", seed=4),
        cases.build_2d(poiseuille(nv=20, ny=8, nm=8), prec),
        cases.build_2d(rt(nv=24, nm=8), prec, rt_variant="cartesian", seed=9),
        cases.build_2d(poiseuille(nv=13, ny=6, nm=6), prec),
    ]


@pytest.mark.parametrize("prec", [capi.FP64, capi.MIXED_FP32])
def test_grouped_batch_equals_cases_stepped_alone(have_gpu, prec):
    """Cases of one storage type run as one grouped launch per kernel type;
    every case must match, bitwise and on every replay, the same case
    stepped alone with the schedule the batch gave it."""
    cs = grouped_cases(prec)
    with capi.Batch(cs) as b:
        print("grouped", b.grouped, "prec", prec, flush=True)
        assert b.grouped
        singles = []
        for c, s in zip(cs, b.solvers):
            x = capi.Solver(c)
            x.set_schedule(*s.schedule())
            singles.append(x)
        for rep, n in enumerate((2, 1, 3)):
            infos = b.advance(n)
            assert [i.steps for i in infos] == [n] * len(cs)
            for k, s in enumerate(singles):
                s.advance(n)
                wa = s.macro(with_q=False)
                wb = b.solvers[k].macro(with_q=False)
                assert np.array_equal(wa, wb), (rep, k, np.max(np.abs(wa - wb)))
                ha, _ = s.get_f()
                hb, _ = b.solvers[k].get_f()
                assert np.array_equal(ha, hb), (rep, k)
        for s in singles:
            s.close()


def test_grouped_batch_parity_and_exits(have_gpu, monkeypatch):
    """A grouped mixed batch against the oracle; with HUGKS_BATCH_GROUPED=0
    the same batch runs case by case and gives the same bits; cases with
    different maxtime stop on their own."""
    cs = grouped_cases(capi.MIXED_FP32)
    with capi.Batch(cs) as b:
        assert b.grouped
        sched = b.solvers[0].schedule()
        assert all(s.schedule() == sched for s in b.solvers)
        assert sched in ((capi.FLUX_CELLSUM, capi.UPDATE_MFMA_FACE),
                         (capi.FLUX_SPLIT, capi.UPDATE_MFMA_FACE))
        b.advance(5)
        wg = [s.macro(with_q=False) for s in b.solvers]
    # case by case with the schedule of the grouped launches, which the
    # batch's graphs capture at create
    monkeypatch.setenv("HUGKS_BATCH_GROUPED", "0")
    monkeypatch.setenv("HUGKS_FLUX_KERNEL",
                       "cellsum" if sched[0] == capi.FLUX_CELLSUM else "split")
    monkeypatch.setenv("HUGKS_UPDATE_KERNEL", "mfma-face")
    with capi.Batch(cs) as b:
        assert not b.grouped
        for s in b.solvers:
            assert s.schedule() == sched
        b.advance(5)
        wl = [s.macro(with_q=False) for s in b.solvers]
    monkeypatch.delenv("HUGKS_BATCH_GROUPED")
    monkeypatch.delenv("HUGKS_FLUX_KERNEL")
    monkeypatch.delenv("HUGKS_UPDATE_KERNEL")
    for c, a, x in zip(cs, wg, wl):
        assert np.array_equal(a, x)
        o = oracle_py.Oracle(c.with_precision(capi.FP64))
        o.step(5)
        err = field_errors(a, o.get()[0])
        print("nv", c.desc.nv, "err", err, flush=True)
        assert np.all(err <= 1e-5), err
    ts = [cases.build_2d(poiseuille(nv=10, ny=8, maxtime=t), capi.FP64) for t in (0.02, 0.05, 0.035)]
    with capi.Batch(ts) as b:
        assert b.grouped
        infos = b.advance(0)
    for c, inf in zip(ts, infos):
        with capi.Solver(c) as s:
            alone = s.advance(0)
        assert inf.stop_reason == capi.STOP_MAXTIME
        assert inf.steps == alone.steps and inf.simtime == alone.simtime


def test_small_cases_grouped_beside_a_large_one(have_gpu):
    """A batch with one case too large for the grouped launches: the small
    cases run grouped, the large one with its own launches beside them, and
    every case matches, bitwise and on every replay, the same case stepped
    alone with the schedule the batch gave it."""
    cs = grouped_cases(capi.MIXED_FP32) + [
        cases.build_2d(rt(nx=130, ny=130, nv=12, nm=6), capi.MIXED_FP32, rt_variant="cartesian",
                       seed=5)]
    with capi.Batch(cs) as b:
        print("grouped cases", b.grouped_cases, "of", len(cs), flush=True)
        assert b.grouped_cases == len(cs) - 1 and not b.grouped
        singles = []
        for c, s in zip(cs, b.solvers):
            x = capi.Solver(c)
            x.set_schedule(*s.schedule())
            singles.append(x)
        for rep, n in enumerate((2, 1)):
            infos = b.advance(n)
            assert [i.steps for i in infos] == [n] * len(cs)
            for k, s in enumerate(singles):
                s.advance(n)
                wa = s.macro(with_q=False)
                wb = b.solvers[k].macro(with_q=False)
                assert np.array_equal(wa, wb), (rep, k, np.max(np.abs(wa - wb)))
        for s in singles:
            s.close()


def test_tune_tiles_picks_a_compiled_tile(have_gpu):
    """hugks_tune_tiles times the wavefronts-per-cell tiles of the MFMA node
    update too; what it returns is a schedule and tile the case accepts, and
    the second call comes from the cache."""
    c = cases.build_2d(poiseuille(nv=20, ny=16), capi.MIXED_FP32)
    fk, uk, nw, ms = capi.tune_tiles(c)
    print("tuned", fk, uk, nw, ms, flush=True)
    assert ms > 0
    with capi.Solver(c) as s:
        s.set_schedule(fk, uk)
        if uk in (capi.UPDATE_MFMA_NODE, capi.UPDATE_MFMA_FACE):
            s.set_update_waves(nw)
            assert s.update_waves() == nw
    again = capi.tune_tiles(c)
    assert again[:3] == (fk, uk, nw) and again[3] == -1.0


def test_batch_cases_refuse_schedule_changes(have_gpu):
    """A batch's graphs hold its cases' launches from create: setting the
    schedule or the node-update tile of one of its handles is refused
    (the schedules they report stay those of the graphs)."""
    cs = grouped_cases(capi.MIXED_FP32)
    with capi.Batch(cs) as b:
        s = b.solvers[0]
        before = s.schedule()
        with pytest.raises(capi.HugksError) as ei:
            s.set_schedule(capi.FLUX_SPLIT, capi.UPDATE_LDS)
        assert ei.value.code == capi.E_ARG
        with pytest.raises(capi.HugksError) as ei:
            s.set_update_waves(1)
        assert ei.value.code == capi.E_ARG
        assert s.schedule() == before


def test_grouped_batch_plain_node_update(have_gpu, monkeypatch):
    """The grouped launches with the plain MFMA node update (the per-face
    node-flux kernel instead of the per-cell divergence), asked for with
    HUGKS_UPDATE_KERNEL=mfma-node: each case as stepped alone, bitwise."""
    monkeypatch.setenv("HUGKS_UPDATE_KERNEL", "mfma-node")
    cs = grouped_cases(capi.MIXED_FP32)
    with capi.Batch(cs) as b:
        assert b.grouped
        scheds = [s.schedule() for s in b.solvers]
        assert all(sc[1] == capi.UPDATE_MFMA_NODE for sc in scheds)
        b.advance(9)
        wg = [s.macro(with_q=False) for s in b.solvers]
    for c, sc, a in zip(cs, scheds, wg):
        with capi.Solver(c) as s:
            s.set_schedule(*sc)
            s.advance(9)
            assert np.array_equal(s.macro(with_q=False), a)
