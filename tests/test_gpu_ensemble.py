"""Asynchronous stepping and the ensemble runner on the device: a system's
result is the same stepped alone, asynchronously, or among others."""
import time

import numpy as np
import pytest

import simcore_amd as sc
from simcore_amd._lib import SC_ERR_NOMEM
from tests.helpers import c0_system, long_system, uneven_system

pytestmark = pytest.mark.gpu


def _systems():
    out = []
    p, f = c0_system(num=32, seed=5, noise_seed=101)
    out.append((p, f, 60))
    p, f = uneven_system(seed=31, noise_seed=202)
    out.append((p, f, 45))
    p, f = c0_system(num=64, seed=6, noise_seed=303)
    out.append((p, f, 33))
    p, f = long_system(seed=41, noise_seed=404)
    out.append((p, f, 50))
    return out


def _alone(p, f, n, precision="mixed"):
    with sc.FilamentSystem(p, f, precision=precision, use_graph=True) as s:
        s.step(n)
        return s.sites()[0]


def test_async_steps_match_sync():
    print("case: C0 32 filaments, 15+25 async steps vs 40 sync", flush=True)
    p, f = c0_system(num=32, seed=5)
    ref = _alone(p, f, 40)
    with sc.FilamentSystem(p, f, precision="mixed", use_graph=True) as s:
        s.step_async(15)
        s.step_async(25)
        t0 = time.time()
        while not s.poll():
            assert time.time() - t0 < 60
            time.sleep(1e-4)
        np.testing.assert_array_equal(s.sites()[0], ref)
        assert s.counters()["i_step"] == 40


def test_ensemble_results_same_alone_and_in_mix():
    print("case: 4 systems alone vs ensemble (max_active 4, 16 steps per launch)", flush=True)
    cases = _systems()
    refs = [_alone(p, f, n) for p, f, n in cases]
    with sc.Ensemble(hbm_budget=2 << 30, max_active=4, steps_per_launch=16) as ens:
        ids = [ens.submit(p, f, n, precision="mixed") for p, f, n in cases]
        ens.drain(timeout_s=120)
        for rid, ref in zip(ids, refs):
            assert ens.request(rid)["state"] == "done"
            np.testing.assert_array_equal(ens.result(rid), ref)
        st = ens.stats()
        print(st)
        assert st["completed"] == len(cases) and st["failed"] == 0
        assert st["peak_active"] >= 2
        assert st["latency_p50"] <= st["latency_p99"] <= st["latency_max"]
        assert st["peak_bytes"] <= st["arena_bytes"]


def test_ensemble_admission_by_footprint():
    print("case: budget for one resident system at a time", flush=True)
    cases = _systems()[:3]
    sizes = []
    for p, f, n in cases:
        with sc.FilamentSystem(p, f, precision="mixed") as s:
            sizes.append(s.device_bytes())
    budget = int(1.3 * max(sizes))
    with sc.Ensemble(hbm_budget=budget, max_active=8, steps_per_launch=8) as ens:
        ids = [ens.submit(p, f, n, precision="mixed") for p, f, n in cases]
        ens.drain(timeout_s=120)
        st = ens.stats()
        print(st, sizes)
        assert st["completed"] == 3
        assert st["peak_bytes"] <= budget
        assert st["peak_active"] == 1
        for rid, (p, f, n) in zip(ids, cases):
            np.testing.assert_array_equal(ens.result(rid), _alone(p, f, n))


def test_ensemble_rejects_request_larger_than_budget():
    print("case: request larger than the whole budget", flush=True)
    p, f = c0_system(num=64)
    with sc.Ensemble(hbm_budget=1 << 18, max_active=2) as ens:
        rid = ens.submit(p, f, 10)
        ens.drain(timeout_s=10)
        info = ens.request(rid)
        assert info["status"] == SC_ERR_NOMEM
        with pytest.raises(sc.SimcoreError):
            ens.result(rid)


def test_ensemble_fp64_requests_match_oracle():
    """FP64 requests run through the ensemble among mixed ones: each
    result equals the CPU oracle's positions after the same steps, bit for
    bit (the ensemble runs the same FP64 path as a system alone)."""
    from oracle.oracle_py import OracleSystem
    print("case: 2 FP64 requests among 2 mixed, vs the oracle", flush=True)
    cases = _systems()
    fp64 = [cases[0], cases[1]]
    with sc.Ensemble(hbm_budget=2 << 30, max_active=4, steps_per_launch=16) as ens:
        ids = [ens.submit(p, f, n, precision="fp64") for p, f, n in fp64]
        ens.submit(*cases[2][:2], cases[2][2], precision="mixed")
        ens.submit(*cases[3][:2], cases[3][2], precision="mixed")
        ens.drain(timeout_s=120)
        for rid, (p, f, n) in zip(ids, fp64):
            assert ens.request(rid)["state"] == "done"
            ora = OracleSystem(p, f)
            assert ora.step(n) == 0
            np.testing.assert_array_equal(ens.result(rid), ora.sites()[0])
            ora.close()


def test_ensemble_admits_while_others_run():
    """Requests submitted while earlier ones are resident and stepping are
    admitted next to them (no waiting for the earlier ones to finish), and
    every result equals the system stepped alone."""
    print("case: submissions during a running request", flush=True)
    cases = _systems()
    p0, f0 = c0_system(num=64, seed=9, noise_seed=505)
    long_steps = 600
    refs = [_alone(p, f, n) for p, f, n in cases]
    ref0 = _alone(p0, f0, long_steps)
    with sc.Ensemble(hbm_budget=2 << 30, max_active=8, steps_per_launch=8) as ens:
        first = ens.submit(p0, f0, long_steps, precision="mixed")
        for _ in range(4):
            ens.progress()
        assert ens.request(first)["state"] == "running"
        ids = []
        for p, f, n in cases:
            ids.append(ens.submit(p, f, n, precision="mixed"))
            ens.progress()
        ens.drain(timeout_s=120)
        info0 = ens.request(first)
        for rid, ref in zip(ids, refs):
            info = ens.request(rid)
            assert info["state"] == "done"
            # admitted before the first request completed
            assert info["t_admit"] < info0["t_done"]
            np.testing.assert_array_equal(ens.result(rid), ref)
        np.testing.assert_array_equal(ens.result(first), ref0)
        assert ens.stats()["peak_active"] >= 2


def _batchable():
    """Mixed-precision systems of different sizes and seeds with filaments of
    at most 128 sites (every length class of the packed integrator), odd and
    even step counts."""
    out = []
    for k, (num, n) in enumerate([(16, 64), (32, 50), (48, 41), (64, 80), (24, 33), (40, 72)]):
        p, f = c0_system(num=num, seed=50 + k, noise_seed=600 + k)
        out.append((p, f, n))
    p, f = uneven_system(sites=[3, 4, 7, 8, 9, 15, 16, 17, 31, 32, 33, 47, 63, 64, 65, 100,
                                127, 128], seed=33, noise_seed=707)
    out.append((p, f, 57))
    return out


@pytest.mark.parametrize("batch_size,spl", [(1, 16), (3, 16), (8, 16), (4, 1)])
def test_ensemble_batches_match_alone(batch_size, spl):
    """Systems stepped in batches -- one set of grouped launches per step for
    several systems, members joining and leaving at chunk boundaries, the
    odd remainders on their own streams -- end with the positions of each
    system stepped alone, bit for bit; batch_size 1 steps every system on
    its own stream; one step per chunk still takes batches two steps at a
    time."""
    cases = _batchable()
    print(f"case: {len(cases)} systems, batch_size {batch_size}, steps per chunk {spl}, "
          f"max_active 5", flush=True)
    refs = [_alone(p, f, n) for p, f, n in cases]
    with sc.Ensemble(hbm_budget=2 << 30, max_active=5, steps_per_launch=spl,
                     batch_size=batch_size) as ens:
        ids = [ens.submit(p, f, n, precision="mixed") for p, f, n in cases]
        ens.drain(timeout_s=120)
        st = ens.stats()
        print(st)
        for rid, ref in zip(ids, refs):
            assert ens.request(rid)["state"] == "done"
            np.testing.assert_array_equal(ens.result(rid), ref)
        total = sum(n for _, _, n in cases)
        assert st["completed"] == len(cases) and st["failed"] == 0
        assert st["batched_steps"] + st["solo_steps"] == total
        if batch_size == 1:
            assert st["batched_steps"] == 0
        else:
            assert st["batched_steps"] >= total - 2 * len(cases)


def test_ensemble_batch_with_rebuilds_and_spatial_order():
    """Spatially ordered mixed systems whose candidate lists are rebuilt
    inside batched chunks (the grouped rebuild launches), a system with
    filaments of every length up to 256 sites (its 129-256-site filaments in
    the first waves of the grouped integrator launch), and an FP64 request,
    which cannot join a batch and runs on its own stream: results equal each
    system alone, rebuild counts included."""
    print("case: spatial order, rebuilds inside batches, long filaments, one FP64 solo",
          flush=True)
    cases = [(p, f, 120, "mixed") for p, f, _ in _batchable()[:4]]
    p, f = long_system(seed=41, noise_seed=404)
    cases.append((p, f, 40, "mixed"))
    p, f = c0_system(num=32, seed=8, noise_seed=808)
    cases.append((p, f, 30, "fp64"))
    refs, rebuilds = [], []
    for p, f, n, prec in cases:
        with sc.FilamentSystem(p, f, precision=prec, use_graph=True,
                               spatial_order=prec == "mixed") as s:
            s.step(n)
            refs.append(s.sites()[0])
            rebuilds.append(s.counters()["n_rebuilds"])
    assert min(rebuilds[:4]) >= 3
    with sc.Ensemble(hbm_budget=2 << 30, max_active=8, steps_per_launch=32,
                     batch_size=8) as ens:
        ids = [ens.submit(p, f, n, precision=prec, spatial_order=prec == "mixed")
               for p, f, n, prec in cases]
        ens.drain(timeout_s=120)
        st = ens.stats()
        print(st)
        for rid, ref in zip(ids, refs):
            np.testing.assert_array_equal(ens.result(rid), ref)
        assert st["batched_steps"] == 4 * 120 + 40
        assert st["solo_steps"] == 30


def test_ensemble_batch_member_grows_its_candidate_capacity():
    """A batch member whose candidate capacity is its exact load-time count
    outgrows it at a rebuild inside a batched chunk: its kernels stop at the
    overflow while the other members go on, it grows its buffers and re-runs
    the rest of the chunk on its own stream, then rejoins the batches.  Every
    result equals the system stepped alone with room to spare."""
    import ctypes
    steps = 240
    # a system whose list grows early: 724 candidates at load, 738 at the
    # rebuild of step 17 (CPU oracle)
    p0, f0 = c0_system(num=64, seed=53, noise_seed=653)
    cases = [(p0, f0, steps)] + [(p, f, steps) for p, f, _ in _batchable()[:3]]
    print(f"case: batch member with exact capacity, {f0.nb} bonds, {steps} steps", flush=True)
    with sc.FilamentSystem(p0, f0, precision="mixed", use_graph=True) as s:
        n0 = s.counters()["n_candidates"]
    tight = type(p0)()
    ctypes.pointer(tight)[0] = p0
    tight.capacity_factor = -float(n0)  # the ensemble takes the parameters as given
    with sc.FilamentSystem(p0, f0, precision="mixed", use_graph=True,
                           capacity_factor=-float(n0)) as s:
        b0 = s.device_bytes()
        s.step(steps)
        assert s.device_bytes() > b0, "the capacity never overflowed in this run"
        grown_alone = s.sites()[0]
    refs = [_alone(p, f, n) for p, f, n in cases]
    np.testing.assert_array_equal(grown_alone, refs[0])
    with sc.Ensemble(hbm_budget=2 << 30, max_active=8, steps_per_launch=16,
                     batch_size=8) as ens:
        ids = [ens.submit(tight, f0, steps, precision="mixed")]
        ids += [ens.submit(p, f, n, precision="mixed") for p, f, n in cases[1:]]
        ens.drain(timeout_s=120)
        st = ens.stats()
        print(st)
        for rid, ref in zip(ids, refs):
            assert ens.request(rid)["state"] == "done"
            np.testing.assert_array_equal(ens.result(rid), ref)
        assert ens.request(ids[0])["device_bytes"] > b0
        assert st["batched_steps"] >= 3 * steps
