"""The rest of mdc's step loop on the device (main.c:44-60): calc_msd
(msd.c:20-81), control_temp (pt.c:85-112) and the moment_correction
cadence (rv.c:88-116), inside the batched step and through the drop-in,
against the oracle (itself pinned bit for bit to the compiled mdc loop)."""
import ctypes as C

import numpy as np
import pytest

import gpuhelp as H
import mdgpu
import mdref

pytestmark = pytest.mark.gpu

# thermostat every 6 steps until step 30 (scalings at 6, 12, 18, 24),
# moment correction every 16 steps, MSD every step
LOOP = dict(t_control_step=6, temp=1400.0, t_until=30, moment_every=16, msd=True)


def _params(o):
    b = o.b
    return mdgpu.box_params(b.nx, b.ny, b.nz, b.Ax, b.a1, b.hm, b.radius, b.delta_time_fs)


def _arm(bt, k=0):
    bt.set_observe(k, msd=True, t_control_step=LOOP["t_control_step"], temp_K=LOOP["temp"],
                   t_control_until=LOOP["t_until"], moment_every=LOOP["moment_every"])


def _start(o, seed):
    r = mdref.jitter_positions(o.r, o.L, 0.1, seed)
    v = mdref.maxwell_velocities(o.masses(), 900.0, seed)
    return r, v


def test_batch_step_loop_matches_oracle(gpu):
    """40 graph-replayed steps of an armed FP64 box against the oracle's
    main.c loop from the same state.  The batch sums forces and
    reductions in another order than mdc, so the trajectories agree to
    rounding growth, not bit for bit."""
    o = H.cfg1_oracle()
    r, v = _start(o, 51)
    o.set_state(r=r, v=v, f=np.zeros_like(r))
    bt = mdgpu.Batch([_params(o)], mdgpu.FP64)
    bt.set_state(0, r=r, v=v)
    _arm(bt)
    o.newton(1, 40, **LOOP)
    bt.step(40)
    st = bt.state(0)
    assert bt.scalars(0)["step"] == 40
    assert H.rel_err(st["r"], o.r) <= 1e-12
    assert H.rel_err(st["v"], o.v) <= 1e-10
    ob = bt.observed(0)
    assert ob["scalings"] == o.obs.scalings == 4
    assert np.allclose(ob["msd"][:2], o.msd_values(), rtol=1e-9, atol=0)
    old, d = bt.msd_state(0)
    assert H.rel_err(old, o.msd_old) <= 1e-12
    assert H.rel_err(d, o.msd_d) <= 1e-9
    # the thermostat acted: without it the trajectory is another one
    free = mdgpu.Batch([_params(o)], mdgpu.FP64)
    free.set_state(0, r=r, v=v)
    free.step(40)
    assert H.rel_err(free.state(0)["v"], st["v"]) > 1e-3
    free.close()
    bt.close()


def test_armed_boxes_step_identically_alone_in_a_batch_and_eagerly(gpu):
    """Per-box determinism holds with the observables armed: two armed
    boxes and an unarmed one step to the same bits in a batch (graph
    replay) as alone, and an armed box steps to the same bits in single
    eager steps as in graph replays."""
    os_ = [H.cfg1_oracle(n=n) for n in (3, 2, 4)]
    ps = [_params(o) for o in os_]
    starts = [_start(o, 60 + k) for k, o in enumerate(os_)]
    armed = (0, 2)
    mix = mdgpu.Batch(ps, mdgpu.FP64)
    for k, (r, v) in enumerate(starts):
        mix.set_state(k, r=r, v=v)
    for k in armed:
        _arm(mix, k)
    mix.step(24)
    for k, p in enumerate(ps):
        for eager in ((False, True) if k == 0 else (False,)):
            lone = mdgpu.Batch([p], mdgpu.FP64)
            lone.set_state(0, r=starts[k][0], v=starts[k][1])
            if k in armed:
                _arm(lone, 0)
            if eager:
                for _ in range(24):
                    lone.step(1)
            else:
                lone.step(24)
            a, b = lone.state(0), mix.state(k)
            for key in ("r", "v", "f"):
                assert np.array_equal(a[key], b[key]), (k, eager, key)
            if k in armed:
                oa, ob = lone.observed(0), mix.observed(k)
                assert np.array_equal(oa["msd"], ob["msd"]) and oa["scalings"] == ob["scalings"]
            lone.close()
    assert mix.observed(1)["scalings"] == 0
    mix.close()


def test_mixed_precision_box_runs_the_same_loop(gpu):
    o = H.cfg1_oracle()
    r, v = _start(o, 71)
    out = {}
    for prec in (mdgpu.FP64, mdgpu.MIXED):
        bt = mdgpu.Batch([_params(o)], prec)
        bt.set_state(0, r=r, v=v)
        _arm(bt)
        bt.step(32)
        out[prec] = bt.observed(0)
        bt.close()
    assert out[mdgpu.MIXED]["scalings"] == out[mdgpu.FP64]["scalings"] == 4
    assert np.allclose(out[mdgpu.MIXED]["msd"][:2], out[mdgpu.FP64]["msd"][:2], rtol=1e-3)


def test_dropin_calc_msd_and_control_temp(gpu):
    """mdc's loop through the drop-in routines on mdc's globals (calc_msd
    and control_temp included) against the oracle's, step by step."""
    o = H.cfg1_oracle()
    d = H.DropIn(o)
    msd = mdgpu.dropin_msd()
    n = o.N
    arrays = np.zeros((6, n))
    value = np.zeros(2)
    number = np.zeros(2, dtype=np.int32)
    dp = C.POINTER(C.c_double)
    try:
        for k, name in enumerate(("rx_old", "ry_old", "rz_old", "dx", "dy", "dz")):
            setattr(msd, name, arrays[k].ctypes.data_as(dp))
        msd.value = value.ctypes.data_as(dp)
        msd.number_of_ion = number.ctypes.data_as(C.POINTER(C.c_int))
        d.sys.e2t = o.b.e2t
        r, v = _start(o, 81)
        o.set_state(r=r, v=v, f=np.zeros_like(r))
        d.set_state(r=r, v=v, f=np.zeros_like(r))
        o.obs.msd_on = 1
        for step in range(1, 21):
            o.b.step = step
            d.sys.step = step
            for name in ("next_rv_verlet", "calc_kin", "calc_press", "calc_msd"):
                o.call(name) if name != "calc_msd" else o.calc_msd()
                d.call(name)
            o.control_temp(5, 1300.0)
            mdgpu.dropin_control_temp(5, 1300.0)
            if step == 1:
                assert np.array_equal(arrays[:3], o.msd_old) and not arrays[3:].any()
        assert H.rel_err(d.r, o.r) <= 1e-12
        assert H.rel_err(d.v, o.v) <= 1e-10
        assert H.rel_err(arrays[:3], o.msd_old) <= 1e-12
        assert H.rel_err(arrays[3:], o.msd_d) <= 1e-9
        assert np.allclose(value, o.msd_values(), rtol=1e-9, atol=0)
        assert list(number) == list(o.obs.number_of_ion[:2])
        assert msd.Lx2 == o.obs.Lx2
    finally:
        for name in ("rx_old", "ry_old", "rz_old", "dx", "dy", "dz", "value", "number_of_ion"):
            setattr(msd, name, None)
        d.release()
