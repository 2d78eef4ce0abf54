"""The configurations' own workloads on the GPU, at their real sizes.

Each per box reproduces mdc's step (mdc/program/rv.c:12-31: calc_foc, then
next_v and next_r), checked three ways: the FP64 path against the oracle
(mdc's arithmetic restated, pinned to the compiled reference) at 1e-12,
the mixed path against FP64 at 1e-5, and a box's bits alone against its
bits inside the full mix.

  cfg4  the 256-box batch of workloads.batch256 (64..4096 ions, NaCl/KCl)
  cfg5  a serving stream of NaCl jobs of 216..32768 ions
  cfg2  the 4096-ion NaCl melt (workloads.melt_snapshot), 5000 FP64 NVE
        steps with cell lists, sampled steps against the oracle
  cfg3  the 32768-ion KCl box, mixed against FP64 from the same seed over
        a stretch of NVE steps
"""
import numpy as np
import pytest

import gpuhelp as H
import mdgpu
import mdref
import workloads as W

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-5


def _table(p):
    q = mdgpu.BoxParams.from_buffer_copy(p)
    q.real_mode = mdgpu.TABLE
    return q


def _batch(boxes, states, prec):
    b = mdgpu.Batch(boxes, prec)
    for k, (r, v, sp) in enumerate(states):
        b.set_state(k, r=r, v=v, species=sp)
    return b


@pytest.fixture(scope="module")
def batch256():
    """cfg4's boxes, their start (the lattice with Maxwell velocities, as
    bench.py starts them) and their state after 16 mixed steps of the full
    batch (two 8-step graph replays)."""
    boxes, temps, salts = W.batch256()
    start = [W.initial_state(p, t, seed=1000 + k) for k, (p, t) in enumerate(zip(boxes, temps))]
    mix = _batch(boxes, start, mdgpu.MIXED)
    mix.step(16)
    now, mixed16 = [], []
    for k in range(len(boxes)):
        st = mix.state(k)
        now.append((st["r"], st["v"], start[k][2]))
        mixed16.append(st)
    mix.close()
    return boxes, salts, start, now, mixed16


def test_batch256_mixed_within_1e5_of_fp64_on_every_box(gpu, batch256):
    boxes, _, _, now, _ = batch256
    f = {}
    for prec in (mdgpu.MIXED, mdgpu.FP64):
        b = _batch(boxes, now, prec)
        b.forces()
        f[prec] = [b.state(k)["f"] for k in range(len(boxes))]
        b.close()
    worst = 0.0
    for k, (f32, f64) in enumerate(zip(f[mdgpu.MIXED], f[mdgpu.FP64])):
        e, erms = H.rel_err(f32, f64), H.rms_rel_err(f32, f64)
        assert e <= MIXED_TOL and erms <= MIXED_TOL, (k, e, erms)
        worst = max(worst, e)
    print(f"batch256: worst max-norm relative force error {worst:.2e}")


def test_batch256_fp64_matches_the_oracle_every_16th_box(gpu, batch256):
    """Every 16th box, in FP64 and in both real-space modes (mdc's table for
    boxes 0, 32, ..., the analytic pair function for 16, 48, ...), against
    the oracle on the same positions: forces, pot and virial to 1e-12."""
    boxes, salts, _, now, _ = batch256
    sel = list(range(0, len(boxes), 16))
    for mode in (mdgpu.TABLE, mdgpu.ANALYTIC):
        ks = [k for k in sel if (k // 16) % 2 == mode]
        ps = [_table(boxes[k]) if mode == mdgpu.TABLE else boxes[k] for k in ks]
        b = _batch(ps, [now[k] for k in ks], mdgpu.FP64)
        b.forces()
        for j, k in enumerate(ks):
            p = boxes[k]
            o = mdref.Oracle(p.nx, p.ny, p.nz, p.ax, p.alpha, p.hm, p.rc, p.dt_fs,
                             salt=salts[k], real_mode=mode)
            r = now[k][0]
            assert np.array_equal(o.ion, now[k][2])
            o.set_state(r=r, f=np.zeros_like(r))
            o.calc_foc()
            g, s = b.state(j)["f"], b.scalars(j)
            assert H.rel_err(g, o.f) <= 1e-12, (k, mode)
            assert s["pot"] == pytest.approx(o.b.pot, rel=1e-12), (k, mode)
            for a, ref in zip(s["vir"], (o.b.virX, o.b.virY, o.b.virZ)):
                assert a == pytest.approx(ref, rel=1e-12, abs=1e-12 * abs(o.b.pot)), (k, mode)
        b.close()


def _energy_slope(boxes, states, prec, steps, every=100):
    """Least-squares slope of each box's consistent total energy (kin + real
    + recip/4 + self) per ion [1e-18 J] over `steps` NVE steps, scaled to
    1e4 steps, from the 20th step on (as bench.py's drift leg)."""
    b = _batch(boxes, states, prec)
    b.step(20)
    es = []
    for s in range(0, steps + 1, every):
        b.sync()
        es.append([b.scalars(k)["etot"] for k in range(len(boxes))])
        if s < steps:
            b.step(every)
    n = np.array([b.n_ions(k) for k in range(len(boxes))], dtype=float)
    b.close()
    es = np.array(es)
    t = np.arange(es.shape[0]) * float(every)
    return np.polyfit(t, es, 1)[0] / n * 1e4


def test_batch256_nve_drift_over_1e4_steps(gpu, batch256):
    """cfg4's NVE drift target, <= 1e-5 per ion per 1e4 steps, over 1e4
    steps at dt = 2 fs (mdc's velocity Verlet, rv.c:33-66) on every 32nd
    box of the batch, mixed and FP64 from the same lattice start.  (A
    2000-step window from the lattice reads about 1e-5: the start's
    relaxation, not a drift, dominates it.)"""
    boxes, _, start, _, _ = batch256
    sel = list(range(0, len(boxes), 32))
    sub = [boxes[k] for k in sel]
    sst = [start[k] for k in sel]
    d32 = _energy_slope(sub, sst, mdgpu.MIXED, 10000)
    d64 = _energy_slope(sub, sst, mdgpu.FP64, 10000)
    assert np.abs(d32).max() <= 1e-5, d32
    assert np.abs(d64).max() <= 1e-5, d64


def test_batch256_boxes_alone_equal_the_full_mix(gpu, batch256):
    """After 16 steps (two graph replays) of all 256 boxes, the smallest,
    the largest and two middle boxes stepped alone hold the same bits."""
    boxes, _, start, _, mixed16 = batch256
    n = np.array([8 * p.nx * p.ny * p.nz for p in boxes])
    order = np.argsort(n, kind="stable")
    for k in (order[0], order[len(order) // 3], order[2 * len(order) // 3], order[-1]):
        lone = _batch([boxes[k]], [start[k]], mdgpu.MIXED)
        lone.step(16)
        st = lone.state(0)
        lone.close()
        for key in ("r", "v", "f", "f0"):
            assert np.array_equal(st[key], mixed16[k][key]), (int(k), key)


def _serve_cfg(**kw):
    c = mdgpu.ServeConfig()
    c.hbm_budget = kw.get("budget", 16 << 30)
    c.max_slots = 64
    c.precision = mdgpu.MIXED
    c.real_mode = mdgpu.ANALYTIC
    c.eps, c.rc_max, c.dt_fs = 1e-6, 10.0, 2.0
    c.one_at_a_time = kw.get("one", 0)
    c.realtime = kw.get("realtime", 1)
    c.steps_per_tick = 8
    return c


def test_stream_216_to_32768_ions_alone_equals_mix_and_budget_binds(gpu):
    """cfg5's stream (NaCl, 216..32768 ions log-uniform, 300..2000 K), job
    lengths scaled to 1/50 of 100..5000 steps, one job set to the largest
    size (16^3 cells, 32768 ions).  Every job's result is bit-identical
    served in the mix and served alone; with a budget that holds the big
    job but not the big job beside all the others, admission waits and the
    results stay the same."""
    jobs = W.stream(seed=11, njobs=8, rate_hz=50.0, n_min=216, n_max=32768, steps_scale=0.02)
    jobs[2].nx = jobs[2].ny = jobs[2].nz = 16
    sizes = [8 * j.nx * j.ny * j.nz for j in jobs]
    assert min(sizes) >= 216 and max(sizes) == 32768
    mix, smix = mdgpu.serve(jobs, _serve_cfg())
    alone, salone = mdgpu.serve(jobs, _serve_cfg(one=1))
    assert smix.max_resident > 1 and salone.max_resident == 1
    for k, (a, b) in enumerate(zip(mix, alone)):
        assert a.status == mdgpu.OK and b.status == mdgpu.OK, k
        assert a.final_checksum == b.final_checksum, k
        assert a.etot_first == b.etot_first and a.etot_last == b.etot_last, k
    total = sum(s * j.steps for s, j in zip(sizes, jobs))
    assert smix.ion_steps == pytest.approx(total)
    assert 0 < smix.p50_latency_s <= smix.p99_latency_s
    # a budget of the big job plus a third of it: the big job runs with
    # only some of the others beside it
    big = mdgpu.Batch([mdgpu.serve_job_params(jobs[2], _serve_cfg())], mdgpu.MIXED)
    need = big.box_bytes(0)
    big.close()
    tight, st = mdgpu.serve(jobs, _serve_cfg(budget=int(1.34 * need), realtime=0))
    assert st.peak_bytes <= int(1.34 * need) and st.rejected == 0
    for a, b in zip(tight, alone):
        assert a.status == mdgpu.OK and a.final_checksum == b.final_checksum


@pytest.fixture(scope="module")
def melt():
    p = W.nacl4096(real_mode=mdgpu.TABLE)
    r, v, sp, msd = W.melt_snapshot(p)
    return p, r, v, sp, msd


def test_melt_snapshot_is_a_liquid(gpu, melt):
    """The cfg2 start is molten: over the 2 ps settle stage at 1500 K both
    species moved several A^2 on average (a rock-salt crystal at this
    temperature vibrates within about 0.5 A^2)."""
    _, _, _, _, msd = melt
    assert min(msd) > 2.0, msd


def test_nacl4096_melt_fp64_trajectory_matches_the_oracle(gpu, melt):
    """cfg2: 5000 FP64 NVE steps of the 4096-ion melt (mdc's table, cell
    lists, rc = 10 A, Ewald at 1e-6).  At steps 0, 2500 and 5000 the state
    is read, one more step is taken, and that step's forces, pot and
    virial -- evaluated on the read positions -- are compared element-wise
    with the oracle's calc_foc on the same positions (real.c:33-83,
    rcprcl.c:32-95) to 1e-12."""
    p, r, v, sp, _ = melt
    b = mdgpu.Batch([p], mdgpu.FP64)
    b.set_state(0, r=r, v=v, species=sp)
    o = mdref.Oracle(8, 8, 8, p.ax, p.alpha, p.hm, p.rc, p.dt_fs, real_mode=0)
    assert np.array_equal(o.ion, sp)
    done, e = 0, []
    for target in (0, 2500, 5000):
        b.step(target - done)
        done = target
        rs = b.state(0)["r"]
        b.step(1)
        done += 1
        g, s = b.state(0)["f"], b.scalars(0)
        e.append(s["etot"])
        o.set_state(r=rs, f=np.zeros_like(rs))
        o.calc_foc()
        assert H.rel_err(g, o.f) <= 1e-12, target
        assert s["pot"] == pytest.approx(o.b.pot, rel=1e-12), target
        for a, ref in zip(s["vir"], (o.b.virX, o.b.virY, o.b.virZ)):
            assert a == pytest.approx(ref, rel=1e-12, abs=1e-12 * abs(o.b.pot)), target
    b.close()
    # NVE: the consistent energy stays put over the 5000 steps
    assert abs(e[-1] - e[0]) / 4096 < 1e-4, e


def test_kcl32768_mixed_tracks_fp64_from_the_same_seed(gpu):
    """cfg3: KCl 32768 ions at 1300 K from the same seeded start, mixed and
    FP64 stepped 200 NVE steps each: the forces at the start agree to 1e-5,
    and the consistent energies stay within 1e-6 per ion of each other at
    every sampled step (the trajectories part only at the FP32 rounding
    level over this stretch)."""
    p = W.kcl32768()
    r, v, sp = W.initial_state(p, 1300.0, seed=33, jitter=0.2)
    out = {}
    for prec in (mdgpu.FP64, mdgpu.MIXED):
        b = mdgpu.Batch([p], prec)
        b.set_state(0, r=r, v=v, species=sp)
        b.forces()
        f0 = b.state(0)["f"]
        b.set_state(0, r=r, v=v)
        es = []
        for _ in range(4):
            b.step(50)
            es.append(b.scalars(0)["etot"])
        out[prec] = (f0, np.array(es))
        b.close()
    (f64, e64), (f32, e32) = out[mdgpu.FP64], out[mdgpu.MIXED]
    assert H.rel_err(f32, f64) <= MIXED_TOL
    assert np.abs(e32 - e64).max() / 32768 <= 1e-6, (e32, e64)
