"""Pins the oracle's calc_msd (msd.c:20-81), control_temp (pt.c:85-112)
and the step loop around them (main.c:26-63): against the compiled mdc
sources where they were built, and against closed forms that a slip in
the fold, the averaging or the order of the loop would fail."""
import numpy as np
import pytest

import mdref

A = 5.64


def box(n=3):
    L = n * A
    return mdref.Oracle(n, n, n, A, 5.6 / L, 27, L / 2, 2.0)


def test_msd_of_a_uniform_drift_across_the_box_edge():
    """Every ion moves by d per sample, wrapped into the box (crossing the
    edges many times): the summed minimum-image displacement after sample
    s is -(s - 1) d for every ion (mdc sums r(t-1) - r(t), msd.c:54-56),
    so MSD = (s - 1)^2 |d|^2 per species."""
    o = box()
    L = o.L[:, None]
    d = np.array([[1.7], [-2.3], [2.9]])
    r0 = o.r.copy()
    for s in range(1, 30):
        o.set_state(r=np.mod(r0 + (s - 1) * d, L))
        o.b.step = s
        o.calc_msd()
        if s == 1:
            assert np.array_equal(o.msd_old, o.r) and not o.msd_d.any()
            continue
        want = (s - 1) ** 2 * float((d ** 2).sum())
        assert np.allclose(o.msd_values(), want, rtol=1e-12, atol=0), s
        assert list(o.obs.number_of_ion[:2]) == [o.N // 2, o.N // 2]
        assert np.allclose(o.msd_d, -(s - 1) * d, rtol=0, atol=1e-12)


def test_msd_fold_is_strict_at_half_box():
    """A displacement of exactly +L/2 is not folded (the tests are strict
    > and <), one just above it is."""
    o = box()
    Lx = o.L[0]
    r0 = o.r.copy()
    o.b.step = 1
    o.calc_msd()
    r1 = r0.copy()
    r1[0, 0] = r0[0, 0] - Lx / 2           # old - new = +L/2: kept
    r1[0, 1] = r0[0, 1] - Lx / 2 - 0.25    # old - new > L/2: folded
    r1[0] = np.where(r1[0] < 0, r1[0] + Lx, r1[0])
    o.set_state(r=r1)
    o.b.step = 2
    o.calc_msd()
    dd0 = r0[0, 0] - r1[0, 0]
    assert o.msd_d[0, 0] == (dd0 if dd0 <= Lx / 2 else dd0 - Lx)
    dd1 = r0[0, 1] - r1[0, 1]
    want1 = dd1 - Lx if dd1 > Lx / 2 else (dd1 + Lx if dd1 < -Lx / 2 else dd1)
    assert o.msd_d[0, 1] == want1
    assert abs(o.msd_d[0, 1]) < Lx / 2


def test_control_temp_hits_the_target_kinetic_energy():
    """With the same kinetic energy K on each of the d_step samples, the
    velocities are scaled once, by sqrt(target / K), at the multiple of
    d_step; the kinetic energy afterwards is the target."""
    o = box()
    v = mdref.maxwell_velocities(o.masses(), 900.0, 4)
    o.set_state(v=v)
    o.call("calc_kin")
    K = o.b.kin
    temp, d = 1500.0, 5
    for s in range(1, d + 1):
        o.b.step = s
        o.b.kin = K
        scaled = o.control_temp(d, temp)
        assert scaled == (s == d)
        if s < d:
            assert np.array_equal(o.v, v)
    target = temp / o.b.e2t
    o.call("calc_kin")
    assert o.b.kin == pytest.approx(target, rel=1e-13)
    assert o.obs.count == 0 and o.obs.oldk == 0.0
    assert np.allclose(o.v / v, np.sqrt(target / K), rtol=1e-15, atol=0)


def test_control_temp_averages_over_the_period():
    """The scale uses the mean of the period's kinetic energies."""
    o = box()
    v = mdref.maxwell_velocities(o.masses(), 900.0, 5)
    o.set_state(v=v)
    ks = [2.0, 3.0, 7.0]
    for s, k in enumerate(ks, start=1):
        o.b.step = s
        o.b.kin = k
        o.control_temp(3, 1000.0)
    assert np.allclose(o.v, v * np.sqrt((1000.0 / o.b.e2t) / np.mean(ks)), rtol=1e-15, atol=0)


@pytest.mark.skipif(not mdref.ref_available(), reason="oracle/_ref not built (no reference checkout)")
def test_newton_loop_matches_compiled_reference():
    """mdc's own step loop (main.c:26-63 without I/O) with calc_msd every
    step, control_temp every 7 steps until step 45 and moment_correction
    every 20: state, MSD arrays and values bit-identical over 60 steps."""
    L = 3 * A
    o = box()
    R = mdref.Reference(3, 3, 3, A, 5.6 / L, 27, L / 2, 2.0)
    R.reset_control_temp()
    r = mdref.jitter_positions(o.r, o.L, 0.15, 31)
    v = mdref.maxwell_velocities(o.masses(), 1100.0, 31)
    o.set_state(r=r, v=v, f=np.zeros_like(r))
    R.set_state(r=r, v=v, f=np.zeros_like(r))
    kw = dict(t_control_step=7, temp=1600.0, t_until=45, moment_every=20, msd=True)
    o.newton(1, 60, **kw)
    R.newton(1, 60, **kw)
    st = R.state()
    for k in ("r", "v", "f"):
        assert np.array_equal(st[k], getattr(o, k)), k
    old, d, val, num = R.msd()
    assert np.array_equal(old, o.msd_old) and np.array_equal(d, o.msd_d)
    assert np.array_equal(val, o.msd_values())
    assert list(num) == list(o.obs.number_of_ion[:2])
    assert o.obs.scalings == 6           # steps 7, 14, ..., 42
    sr, so = R.scalars(), o.scalars()
    for k in sr:
        assert sr[k] == so[k], k
