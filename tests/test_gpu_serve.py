"""Serving loop: continuous batching under an HBM budget.

Each job's result must be the same alone and in the mix: the engine's
per-box arithmetic does not depend on the other resident boxes, so the
final positions (checksum) and energies agree exactly."""
import numpy as np
import pytest

import mdgpu
import workloads as W

pytestmark = pytest.mark.gpu


def _cfg(**kw):
    c = mdgpu.ServeConfig()
    c.hbm_budget = kw.get("budget", 8 << 30)
    c.max_slots = kw.get("slots", 64)
    c.precision = kw.get("precision", mdgpu.MIXED)
    c.real_mode = mdgpu.ANALYTIC
    c.eps = 1e-6
    c.rc_max = 10.0
    c.dt_fs = 2.0
    c.one_at_a_time = kw.get("one", 0)
    c.realtime = kw.get("realtime", 1)
    c.steps_per_tick = kw.get("tick", 8)
    return c


def _jobs(n=10, seed=3):
    return W.stream(seed=seed, njobs=n, rate_hz=200.0, n_min=64, n_max=1728, steps_min=10,
                    steps_max=60)


def test_results_identical_alone_and_in_mix(gpu):
    jobs = _jobs()
    mix, smix = mdgpu.serve(jobs, _cfg())
    alone, salone = mdgpu.serve(jobs, _cfg(one=1))
    assert smix.max_resident > 1 and salone.max_resident == 1
    for a, b in zip(mix, alone):
        assert a.status == mdgpu.OK and b.status == mdgpu.OK
        assert a.final_checksum == b.final_checksum
        assert a.etot_last == b.etot_last and a.etot_first == b.etot_first
    total = sum(8 * j.nx * j.ny * j.nz * j.steps for j in jobs)
    assert smix.ion_steps == pytest.approx(total) and salone.ion_steps == pytest.approx(total)
    assert 0 < smix.p50_latency_s <= smix.p99_latency_s


def test_budget_binds_and_oversized_job_is_rejected(gpu):
    jobs = _jobs(n=8, seed=5)
    big = mdgpu.Job()
    big.arrival_s, big.nx, big.ny, big.nz = 0.0, 10, 10, 10
    big.temp_k, big.steps, big.seed = 1000.0, 5, 1
    jobs.append(big)
    # budget that holds a few of the small jobs but not the 8000-ion one
    sizes = []
    for j in jobs[:-1]:
        b = mdgpu.Batch([W.ewald_box(j.nx, j.ny, j.nz, W.A_NACL, "NaCl")], mdgpu.MIXED)
        sizes.append(b.box_bytes(0))
        b.close()
    budget = int(2.5 * max(sizes))
    res, st = mdgpu.serve(jobs, _cfg(budget=budget, realtime=0))
    assert res[-1].status == mdgpu.E_BUDGET and st.rejected == 1
    assert all(r.status == mdgpu.OK for r in res[:-1])
    assert st.peak_bytes <= budget
    assert 1 <= st.max_resident < len(jobs) - 1     # the budget, not the slots, binds
    ref, _ = mdgpu.serve(jobs[:-1], _cfg(realtime=0))
    for a, b in zip(res[:-1], ref):
        assert a.final_checksum == b.final_checksum


def test_serving_energy_is_conserved(gpu):
    jobs = _jobs(n=4, seed=7)
    res, _ = mdgpu.serve(jobs, _cfg(precision=mdgpu.FP64, realtime=0))
    for r, j in zip(res, jobs):
        n = 8 * j.nx * j.ny * j.nz
        # start from the lattice: energy moves only by the O(dt^2) VV fluctuation
        assert abs(r.etot_last - r.etot_first) / n < 5e-4, (j.nx, j.ny, j.nz, j.temp_k)
