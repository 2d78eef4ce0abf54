"""Parity of the HIP path (through the C ABI) with the oracle.

Bar (the plain FP64 path): relative error <= 1e-12 against the oracle on the
same inputs, norm-wise over all force components (max |dF| / max |F|), and
relative on pot and each virial component.  Table-mode real-space forces
from the all-pairs kernel are compared bit for bit.
"""
import numpy as np
import pytest

import gpuhelp as H
import mdgpu
import mdref

pytestmark = pytest.mark.gpu
TOL = 1e-12


def states(o, seeds=(1, 2), amp=0.3):
    for s in seeds:
        yield (mdref.jitter_positions(o.r, o.L, amp, s),
               mdref.maxwell_velocities(o.masses(), 1000.0, s))


def oracle_forces(o, r, real_only=False):
    o.set_state(r=r, f=np.zeros_like(r))
    if real_only:
        o.call("clear_foc")
        o.call("real_space" if o.b.real_mode == 0 else "real_space_analytic")
    else:
        o.calc_foc()
    return o.f.copy(), o.scalars()


def test_dropin_real_space_bit_identical(gpu):
    o = H.cfg1_oracle()
    d = H.DropIn(o)
    try:
        for r, v in states(o, seeds=(1, 2, 3)):
            fo, so = oracle_forces(o, r, real_only=True)
            d.set_state(r=r, v=v, f=np.zeros_like(r))
            d.call("real_space")
            assert np.array_equal(d.f, fo), H.rel_err(d.f, fo)
            assert d.sys.pot == pytest.approx(so["pot"], rel=TOL)
            for k in ("virX", "virY", "virZ"):
                assert getattr(d.sys, k) == pytest.approx(so[k], rel=TOL)
    finally:
        d.release()


def test_dropin_calc_foc_parity(gpu):
    o = H.cfg1_oracle()
    d = H.DropIn(o)
    try:
        for r, v in states(o, seeds=(4, 5)):
            f_prev = np.random.default_rng(0).standard_normal(r.shape)
            fo, so = oracle_forces(o, r)
            d.set_state(r=r, v=v, f=f_prev)
            d.call("calc_foc")
            assert H.rel_err(d.f, fo) <= TOL
            assert np.array_equal(d.f0, f_prev)          # clear_foc: f0 <- f
            assert d.sys.pot == pytest.approx(so["pot"], rel=TOL)
            for k in ("virX", "virY", "virZ"):
                assert getattr(d.sys, k) == pytest.approx(so[k], rel=TOL)
        # the routines one by one, as mdc's calc_foc sequences them (rv.c:12-18)
        r, v = next(states(o, seeds=(6,)))
        fo, so = oracle_forces(o, r)
        d.set_state(r=r, v=v, f=np.zeros_like(r))
        for name in ("clear_foc", "real_space", "reciprocal_space", "reciprocal_space3"):
            d.call(name)
        assert H.rel_err(d.f, fo) <= TOL
        assert d.sys.pot == pytest.approx(so["pot"], rel=TOL)
    finally:
        d.release()


def test_dropin_trajectory_and_observables(gpu):
    o = H.cfg1_oracle()
    d = H.DropIn(o)
    try:
        r, v = next(states(o, seeds=(7,), amp=0.1))
        o.set_state(r=r, v=v, f=np.zeros_like(r))
        d.set_state(r=r, v=v, f=np.zeros_like(r))
        for step in range(1, 11):
            o.b.step = step
            o.call("next_rv_verlet")
            o.call("calc_kin")
            o.call("calc_press")
            d.call("next_rv_verlet")
            d.call("calc_kin")
            d.call("calc_press")
        # 10 steps from identical inputs: differences stay at rounding level
        assert H.rel_err(d.r, o.r) <= 1e-13
        assert H.rel_err(d.v, o.v) <= 1e-11
        s = o.scalars()
        for k in ("kin", "pot", "pres"):
            assert getattr(d.sys, k) == pytest.approx(s[k], rel=1e-11), k
        # kinetic energy and momentum correction on identical velocities
        o.set_state(v=d.v)
        o.call("calc_kin")
        d.call("calc_kin")
        assert d.sys.kin == pytest.approx(o.b.kin, rel=TOL)
        o.call("moment_correction")
        d.call("moment_correction")
        assert H.rel_err(d.v, o.v) <= TOL
        # next_v / next_r alone
        f0 = d.f.copy() * 0.9
        d.set_state(f0=f0)
        o.f0[...] = f0
        o.set_state(r=d.r, v=d.v, f=d.f)
        o.call("next_v")
        d.call("next_v")
        assert np.array_equal(d.v, o.v)
        o.call("next_r")
        d.call("next_r")
        assert np.array_equal(d.r, o.r)
    finally:
        d.release()


def _batch_for(o, precision=mdgpu.FP64, real_mode=mdgpu.TABLE, cells=mdgpu.CELLS_AUTO,
               salt="NaCl"):
    b = o.b
    p = mdgpu.box_params(b.nx, b.ny, b.nz, b.Ax, b.a1, b.hm, b.radius, b.delta_time_fs,
                         real_mode=real_mode, cells=cells, salt=salt)
    return mdgpu.Batch([p], precision)


@pytest.mark.parametrize("real_mode", [mdgpu.TABLE, mdgpu.ANALYTIC])
def test_batch_fp64_cfg1(gpu, real_mode):
    o = H.cfg1_oracle(real_mode=real_mode)
    bt = _batch_for(o, real_mode=real_mode)
    for r, v in states(o, seeds=(11, 12)):
        fo, so = oracle_forces(o, r)
        bt.set_state(0, r=r, v=v)
        bt.forces()
        st, sc = bt.state(0), bt.scalars(0)
        assert H.rel_err(st["f"], fo) <= TOL
        assert sc["pot"] == pytest.approx(so["pot"], rel=TOL)
        assert sc["pot_recip"] == pytest.approx(so["pot_recip"], rel=1e-11)
        for k, i in (("virX", 0), ("virY", 1), ("virZ", 2)):
            assert sc["vir"][i] == pytest.approx(so[k], rel=TOL)


@pytest.mark.parametrize("real_mode", [mdgpu.TABLE, mdgpu.ANALYTIC])
def test_batch_fp64_cells_cfg2_size(gpu, real_mode):
    """NaCl 4096 ions (8^3 cells), rc = 10 A, alpha and hm for 1e-6: cell
    lists on the GPU against the oracle's all-pairs loop."""
    a, hm = mdgpu.ewald_params(10.0, 8 * H.A_NACL, 1e-6)
    o = mdref.Oracle(8, 8, 8, H.A_NACL, a, hm, 10.0, 2.0, real_mode=real_mode)
    bt = _batch_for(o, real_mode=real_mode, cells=mdgpu.CELL_LIST)
    r, v = next(states(o, seeds=(21,), amp=0.4))
    fo, so = oracle_forces(o, r)
    bt.set_state(0, r=r, v=v)
    bt.forces()
    st, sc = bt.state(0), bt.scalars(0)
    assert H.rel_err(st["f"], fo) <= TOL
    assert sc["pot"] == pytest.approx(so["pot"], rel=TOL)
    for k, i in (("virX", 0), ("virY", 1), ("virZ", 2)):
        assert sc["vir"][i] == pytest.approx(so[k], rel=TOL)
    w = bt.work(0)
    assert w["nk"] == mdref.lib().orc_count_kvectors(hm)


def test_batch_kcl(gpu):
    o = H.cfg1_oracle(a=H.A_KCL, salt="KCl", real_mode=mdgpu.ANALYTIC)
    bt = _batch_for(o, real_mode=mdgpu.ANALYTIC, salt="KCl")
    r, v = next(states(o, seeds=(31,)))
    fo, so = oracle_forces(o, r)
    bt.set_state(0, r=r, v=v)
    bt.forces()
    assert H.rel_err(bt.state(0)["f"], fo) <= TOL
    assert bt.scalars(0)["pot"] == pytest.approx(so["pot"], rel=TOL)


def test_batch_trajectory_matches_oracle(gpu):
    o = H.cfg1_oracle()
    bt = _batch_for(o)
    r, v = next(states(o, seeds=(41,), amp=0.1))
    o.set_state(r=r, v=v, f=np.zeros_like(r))
    bt.set_state(0, r=r, v=v)
    o.nve(10)
    bt.step(10)
    st = bt.state(0)
    assert H.rel_err(st["r"], o.r) <= 1e-13
    assert H.rel_err(st["v"], o.v) <= 1e-11
    sc = bt.scalars(0)
    assert sc["step"] == 10
    assert sc["kin"] == pytest.approx(o.b.kin, rel=1e-11)
    assert sc["pot"] == pytest.approx(o.b.pot, rel=1e-11)


def test_dropin_beyond_ordered_size_uses_cells(gpu):
    """Above N = 8192 the drop-in's real space runs on cell lists (partner
    order differs from mdc's, so the forces agree to rounding, not bit for
    bit): NaCl 11^3 cells, 10648 ions, rc = 10 A, against the oracle's
    all-pairs loop."""
    a, hm = mdgpu.ewald_params(10.0, 11 * H.A_NACL, 1e-6)
    o = mdref.Oracle(11, 11, 11, H.A_NACL, a, min(hm, 60), 10.0, 2.0)
    d = H.DropIn(o)
    try:
        r, v = next(states(o, seeds=(61,), amp=0.2))
        fo, so = oracle_forces(o, r)
        d.set_state(r=r, v=v, f=np.zeros_like(r))
        d.call("calc_foc")
        assert o.N == 10648
        assert H.rel_err(d.f, fo) <= TOL
        assert d.sys.pot == pytest.approx(so["pot"], rel=TOL)
        for k in ("virX", "virY", "virZ"):
            assert getattr(d.sys, k) == pytest.approx(so[k], rel=TOL)
    finally:
        d.release()
