"""Edge cases of the boundary on the GPU: Ewald switched off by mdc's
sentinel, pairs inside the table's start and non-finite coordinates
(reported, never faulting), argument errors, and ragged box sizes."""
import numpy as np
import pytest

import gpuhelp as H
import mdgpu
import mdref

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("alpha", [11.0])
def test_ewald_sentinel_turns_reciprocal_space_off(gpu, alpha):
    """alpha > 10 skips both reciprocal terms (mdc/program/rcprcl.c:23,115):
    forces are the real-space forces alone and equal the oracle's with the
    same sentinel.  (The other sentinel, alpha == 0, leaves mdc's table
    non-finite for charged ions: Mori's erfc divides by e^0 - 1 at x = 0,
    ext.c:73-104; the device reproduces that too.)"""
    n, a = 3, H.A_NACL
    L = n * a
    o = mdref.Oracle(n, n, n, a, alpha, 27, L / 2, 2.0)
    r = mdref.jitter_positions(o.r, o.L, 0.2, 3)
    o.set_state(r=r, f=np.zeros_like(r))
    o.calc_foc()
    p = mdgpu.box_params(n, n, n, a, alpha, 27, L / 2, 2.0, real_mode=mdgpu.TABLE)
    for prec, tol in ((mdgpu.FP64, 1e-12), (mdgpu.MIXED, 1e-5)):
        b = mdgpu.Batch([p], prec)
        b.set_state(0, r=r)
        b.forces()
        s = b.scalars(0)
        assert s["pot_recip"] == 0.0 and s["pot_self"] == 0.0
        assert b.work(0)["nk"] == 0
        assert H.rel_err(b.state(0)["f"], o.f) <= tol
        assert s["pot"] == pytest.approx(o.b.pot, rel=tol)
        b.close()


def test_empty_half_sphere_keeps_the_self_term(gpu):
    """hm = 0 with a valid alpha: no k vector (rcprcl.c:32-38), yet
    reciprocal_space3 adds -a3 sum Z^2 (rcprcl.c:108-133).  The handle API's
    calc_foc, in both precisions, and the drop-in's calc_foc and its
    routine-by-routine sequence all agree with the oracle, self term
    included."""
    n, a = 3, H.A_NACL
    L = n * a
    o = H.cfg1_oracle(hm=0)
    r = mdref.jitter_positions(o.r, o.L, 0.2, 31)
    o.set_state(r=r, f=np.zeros_like(r))
    o.calc_foc()
    assert o.b.pot_self != 0.0 and o.b.pot_recip == 0.0
    p = mdgpu.box_params(n, n, n, a, 5.6 / L, 0, L / 2, 2.0, real_mode=mdgpu.TABLE)
    for prec, tol in ((mdgpu.FP64, 1e-12), (mdgpu.MIXED, 1e-5)):
        b = mdgpu.Batch([p], prec)
        b.set_state(0, r=r)
        b.forces()
        s = b.scalars(0)
        assert b.work(0)["nk"] == 0
        assert s["pot_recip"] == 0.0
        assert s["pot_self"] == pytest.approx(o.b.pot_self, rel=1e-14)
        assert s["pot"] == pytest.approx(o.b.pot, rel=tol)
        assert H.rel_err(b.state(0)["f"], o.f) <= tol
        b.close()
    d = H.DropIn(o)
    try:
        for seq in (("calc_foc",), ("clear_foc", "real_space", "reciprocal_space",
                                    "reciprocal_space3")):
            d.set_state(r=r, f=np.zeros_like(r), f0=np.zeros_like(r))
            for name in seq:
                d.call(name)
            assert H.rel_err(d.f, o.f) <= 1e-12, seq
            assert d.sys.pot == pytest.approx(o.b.pot, rel=1e-12), seq
    finally:
        d.release()


def test_pair_inside_table_start_is_reported(gpu):
    """Table mode covers r >= 0.5 A (mdc's bins, init.c:65); a closer pair
    is an error mdc would index out of its table for.  Both precisions
    report MDG_E_RANGE at the next scalar read; the kernels run to the end."""
    p = mdgpu.box_params(2, 2, 2, H.A_NACL, 0.5, 27, 5.0, 2.0, real_mode=mdgpu.TABLE)
    r, sp = mdgpu.rocksalt(2, 2, 2, H.A_NACL)
    r = r.copy()
    r[:, 1] = r[:, 0] + np.array([0.3, 0.0, 0.0])
    for prec in (mdgpu.FP64, mdgpu.MIXED):
        b = mdgpu.Batch([p], prec)
        b.set_state(0, r=r, species=sp)
        b.forces()
        with pytest.raises(mdgpu.MdgError) as e:
            b.scalars(0)
        assert e.value.code == mdgpu.E_RANGE
        b.close()


def test_non_finite_position_is_rejected_or_reported(gpu):
    """A NaN coordinate never reaches the kernels silently: set_state
    rejects it, or the step reports MDG_E_STATE."""
    p = mdgpu.box_params(2, 2, 2, H.A_NACL, 0.5, 27, 5.0, 2.0)
    r, sp = mdgpu.rocksalt(2, 2, 2, H.A_NACL)
    r = r.copy()
    r[1, 5] = np.nan
    b = mdgpu.Batch([p], mdgpu.FP64)
    try:
        b.set_state(0, r=r, species=sp)
    except mdgpu.MdgError as e:
        assert e.code in (mdgpu.E_ARG, mdgpu.E_STATE)
    else:
        b.step(2)
        with pytest.raises(mdgpu.MdgError) as e:
            b.scalars(0)
        assert e.value.code == mdgpu.E_STATE
    b.close()


def test_argument_errors(gpu):
    p = mdgpu.box_params(2, 2, 2, H.A_NACL, 0.5, 27, 5.0, 2.0, real_mode=mdgpu.ANALYTIC)
    q = mdgpu.box_params(2, 2, 2, H.A_NACL, 0.5, 27, 5.0, 2.0, real_mode=mdgpu.TABLE)
    with pytest.raises(mdgpu.MdgError) as e:      # one pair kernel per launch
        mdgpu.Batch([p, q], mdgpu.MIXED)
    assert e.value.code == mdgpu.E_ARG
    b = mdgpu.Batch([p], mdgpu.MIXED)
    with pytest.raises(mdgpu.MdgError) as e:
        b.scalars(1)
    assert e.value.code == mdgpu.E_ARG
    with pytest.raises(mdgpu.MdgError) as e:
        b.step_recorded(10, 3)                     # not a multiple of `every`
    assert e.value.code == mdgpu.E_ARG
    with pytest.raises(mdgpu.MdgError) as e:
        b.set_observe(0, t_control_step=-1)
    assert e.value.code == mdgpu.E_ARG
    b.close()
