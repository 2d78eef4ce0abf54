"""Mixed precision against FP64, batching invariance and graph replay."""
import numpy as np
import pytest

import gpuhelp as H
import mdgpu
import mdref

pytestmark = pytest.mark.gpu

# the mixed path's bound: relative force error <= 1e-5 against FP64
MIXED_TOL = 1e-5


def _params(n, a, eps=1e-6, rc_max=10.0, salt="NaCl", real_mode=mdgpu.ANALYTIC,
            cells=mdgpu.CELLS_AUTO):
    nx, ny, nz = n if isinstance(n, tuple) else (n, n, n)
    L = np.array([nx, ny, nz]) * a
    rc = min(rc_max, 0.5 * L.min())
    alpha, hm = mdgpu.ewald_params(rc, L.max(), eps)
    return mdgpu.box_params(nx, ny, nz, a, alpha, hm, rc, 2.0, real_mode=real_mode,
                            cells=cells, salt=salt)


def _state(p, seed, temp=1500.0, amp=0.3):
    r, sp = mdgpu.rocksalt(p.nx, p.ny, p.nz, (p.ax, p.ay, p.az))
    L = np.array([p.nx * p.ax, p.ny * p.ay, p.nz * p.az])
    r = mdref.jitter_positions(r, L, amp, seed)
    v = mdgpu.maxwell(seed, temp, sp, p.pot)
    return r, v, sp


@pytest.mark.parametrize("n,salt,a", [(3, "NaCl", H.A_NACL), (8, "NaCl", H.A_NACL),
                                      (6, "KCl", H.A_KCL)])
def test_mixed_forces_within_1e5_of_fp64(gpu, n, salt, a):
    p = _params(n, a, salt=salt)
    r, v, sp = _state(p, 5)
    res = {}
    for prec in (mdgpu.FP64, mdgpu.MIXED):
        b = mdgpu.Batch([p], prec)
        b.set_state(0, r=r, v=v, species=sp)
        b.forces()
        res[prec] = (b.state(0)["f"], b.scalars(0))
        b.close()
    f64, s64 = res[mdgpu.FP64]
    f32, s32 = res[mdgpu.MIXED]
    assert H.rel_err(f32, f64) <= MIXED_TOL
    assert H.rms_rel_err(f32, f64) <= MIXED_TOL
    assert s32["etot"] == pytest.approx(s64["etot"], rel=1e-6)
    for i in range(3):
        assert s32["vir"][i] == pytest.approx(s64["vir"][i], rel=1e-5)


def test_batch_members_match_lone_boxes(gpu):
    """Each box steps to identical bits alone and inside a mixed batch."""
    specs = [(2, "NaCl", H.A_NACL), (3, "KCl", H.A_KCL), (5, "NaCl", H.A_NACL),
             ((2, 3, 4), "NaCl", H.A_NACL)]
    for prec in (mdgpu.MIXED, mdgpu.FP64):
        ps = [_params(n, a, salt=s) for n, s, a in specs]
        sts = [_state(p, 100 + k) for k, p in enumerate(ps)]
        mix = mdgpu.Batch(ps, prec)
        for k, (r, v, sp) in enumerate(sts):
            mix.set_state(k, r=r, v=v, species=sp)
        mix.step(20)
        for k, p in enumerate(ps):
            lone = mdgpu.Batch([p], prec)
            r, v, sp = sts[k]
            lone.set_state(0, r=r, v=v, species=sp)
            lone.step(20)
            a_, b_ = lone.state(0), mix.state(k)
            for key in ("r", "v", "f"):
                assert np.array_equal(a_[key], b_[key]), (prec, k, key)
            assert lone.scalars(0)["etot"] == mix.scalars(k)["etot"]
            lone.close()
        mix.close()


def test_concurrent_box_groups_match_lone_boxes(gpu):
    """With 16 or more boxes a step runs two box groups and each box's real
    space and reciprocal chain on separate streams.  Kernels running side by
    side leave each other's data in LDS; every box must still step to the
    same (finite) bits as alone.  Regression: the mixed real space's idle
    lanes once read a partner slot that the window had not staged."""
    specs = [(2, "NaCl", H.A_NACL), (3, "KCl", H.A_KCL), (4, "NaCl", H.A_NACL),
             ((2, 3, 4), "KCl", H.A_KCL), (5, "NaCl", H.A_NACL)] * 4
    ps = [_params(n, a, salt=s) for n, s, a in specs]
    sts = [_state(p, 500 + k) for k, p in enumerate(ps)]
    mix = mdgpu.Batch(ps, mdgpu.MIXED)
    for k, (r, v, sp) in enumerate(sts):
        mix.set_state(k, r=r, v=v, species=sp)
    mix.step(200)
    for k in range(0, len(ps), 3):
        lone = mdgpu.Batch([ps[k]], mdgpu.MIXED)
        r, v, sp = sts[k]
        lone.set_state(0, r=r, v=v, species=sp)
        lone.step(200)
        a_, b_ = lone.state(0), mix.state(k)
        assert np.isfinite(b_["r"]).all() and np.isfinite(b_["f"]).all(), k
        for key in ("r", "v", "f"):
            assert np.array_equal(a_[key], b_[key]), (k, key)
        lone.close()
    mix.close()


def test_tune_schedule_leaves_the_trajectory_unchanged(gpu):
    """mdg_batch_tune_schedule steps the batch under several group splits
    but restores every box's device state: right after it the states are
    the same bits, and 24 steps later (graph replays under the chosen
    split) they equal a batch that never tuned."""
    ps, states = [], []
    for k in range(20):
        n = 2 + k % 4
        p = _params(n, H.A_KCL if k % 3 == 0 else H.A_NACL,
                    salt="KCl" if k % 3 == 0 else "NaCl")
        ps.append(p)
        states.append(_state(p, 300 + k, amp=0.2))
    a, b = (mdgpu.Batch(ps, mdgpu.MIXED) for _ in range(2))
    for x in (a, b):
        for k, (r, v, sp) in enumerate(states):
            x.set_state(k, r=r, v=v, species=sp)
        x.step(16)
    before = [a.state(k) for k in range(len(ps))]
    share = a.tune_schedule(16)
    assert 0.3 <= share <= 0.7  # the tuner's candidate splits
    for k in range(len(ps)):
        st = a.state(k)
        for key in ("r", "v", "f", "f0"):
            assert np.array_equal(st[key], before[k][key]), (k, key)
    for x in (a, b):
        x.step(24)
    for k in range(len(ps)):
        sa, sb = a.state(k), b.state(k)
        for key in ("r", "v", "f", "f0"):
            assert np.array_equal(sa[key], sb[key]), (k, key)
        assert a.scalars(k)["step"] == b.scalars(k)["step"]
    a.close()
    b.close()


def test_profile_times_the_step_schedule(gpu):
    """profile() times every launch of the step as the step schedules it:
    with 16 or more boxes the reciprocal chain runs once per box group, the
    real space and the update once per step; and profiling steps the state
    exactly as step() does."""
    ps = [_params(2, H.A_NACL)] * 16
    r, v, sp = _state(ps[0], 1)
    out = []
    for prof_first in (True, False):
        b = mdgpu.Batch(ps, mdgpu.MIXED)
        for k in range(len(ps)):
            b.set_state(k, r=r, v=v, species=sp)
        if prof_first:
            prof = b.profile(3)
        else:
            b.step(3)
        out.append(b.state(5)["r"])
        b.close()
    assert np.array_equal(out[0], out[1])
    assert prof["real_space"][1] == 1 and prof["force_vv"][1] == 1
    for k in ("phase_tables", "sk_partial", "sk_finalize", "kf_pack", "kforce"):
        assert prof[k][1] == 2, (k, prof[k])
    assert all(ms > 0 for ms, _ in prof.values())


def test_graph_replay_equals_eager_steps(gpu):
    p = _params(4, H.A_NACL)
    r, v, sp = _state(p, 7)
    out = []
    for chunks in ([8, 8], [1] * 16):   # 8-step graphs vs single eager steps
        b = mdgpu.Batch([p], mdgpu.MIXED)
        b.set_state(0, r=r, v=v, species=sp)
        for c in chunks:
            b.step(c)
        out.append(b.state(0))
        assert b.scalars(0)["step"] == 16
        b.close()
    for key in ("r", "v", "f", "f0"):
        assert np.array_equal(out[0][key], out[1][key]), key


def nve_drift(p, prec, r, v, sp, steps=1000, every=10):
    """Least-squares slope of the consistent total energy per ion
    [1e-18 J], scaled to 1e4 steps (the first sample, which still carries
    mdc's f0 = 0 start, is left out), and the peak-to-peak fluctuation."""
    b = mdgpu.Batch([p], prec)
    b.set_state(0, r=r, v=v, species=sp)
    es = []
    for _ in range(steps // every):
        b.step(every)
        es.append(b.scalars(0)["etot"])
    n = b.n_ions(0)
    b.close()
    es = np.array(es[1:]) / n
    t = np.arange(es.size) * every
    return np.polyfit(t, es, 1)[0] * 1e4, np.ptp(es)


@pytest.mark.parametrize("n,salt,a", [(8, "NaCl", H.A_NACL), ((6, 7, 8), "KCl", H.A_KCL),
                                      ((2, 3, 4), "NaCl", H.A_NACL)])
def test_mixed_screen_keeps_every_pair(gpu, n, salt, a):
    """The mixed real space screens candidates on 16-bit periodic
    coordinates before its exact FP64 cutoff test: it must accept exactly
    the FP64 path's pairs, with all pairs and with cell lists, cubic or
    not (rc = L_min / 2 in the 2x3x4 box)."""
    modes = [mdgpu.ALL_PAIRS]
    p0 = _params(n, a, salt=salt)
    nx, ny, nz = n if isinstance(n, tuple) else (n, n, n)
    if min(np.array([nx, ny, nz]) * a) >= 3 * p0.rc:
        modes.append(mdgpu.CELL_LIST)
    r, v, sp = _state(p0, 11, amp=0.5)
    b = mdgpu.Batch([p0], mdgpu.FP64)
    b.set_state(0, r=r, v=v, species=sp)
    b.forces()
    f64, pairs64 = b.state(0)["f"], b.work(0)["pairs"]
    b.close()
    got = []
    for cells in modes:
        p = _params(n, a, salt=salt, cells=cells)
        b = mdgpu.Batch([p], mdgpu.MIXED)
        b.set_state(0, r=r, v=v, species=sp)
        b.forces()
        f32, pairs32 = b.state(0)["f"], b.work(0)["pairs"]
        b.close()
        assert pairs32 == pairs64, (cells, pairs32, pairs64)
        assert H.rel_err(f32, f64) <= MIXED_TOL
        got.append(f32)
    if len(got) == 2:  # same pairs, other summation order
        assert H.rel_err(got[1], got[0]) <= 1e-6


def test_mixed_nve_energy_drift(gpu):
    """NVE (analytic real space, dt = 1 fs, 1000 K): the mixed path's energy
    drift is <= 1e-5 per ion per 1e4 steps and equals the FP64 path's to
    1e-6 (the drift comes from the unshifted cutoff, not from FP32)."""
    p = _params(4, H.A_NACL)
    p.dt_fs = 1.0
    r, v, sp = _state(p, 9, temp=1000.0, amp=0.05)
    d32, fl32 = nve_drift(p, mdgpu.MIXED, r, v, sp)
    d64, fl64 = nve_drift(p, mdgpu.FP64, r, v, sp)
    assert abs(d32) <= 1e-5
    assert abs(d32 - d64) <= 1e-6
    assert fl32 == pytest.approx(fl64, rel=0.1)


def test_autotuned_tilings_keep_parity(gpu):
    """Tuned tilings change only how work is split: FP64 forces agree with
    the built-in tiling's to 1e-12, mixed stays within 1e-5 of FP64, and a
    tuned box computes the same bits alone and in a batch."""
    ps = [_params(3, H.A_NACL), _params(8, H.A_NACL), _params(6, H.A_KCL, salt="KCl"),
          _params((2, 3, 4), H.A_NACL)]
    sts = [_state(p, 300 + k) for k, p in enumerate(ps)]

    def forces(prec):
        b = mdgpu.Batch(ps, prec)
        for k, (r, v, sp) in enumerate(sts):
            b.set_state(k, r=r, v=v, species=sp)
        b.forces()
        out = [b.state(k)["f"] for k in range(len(ps))]
        b.close()
        return out

    mdgpu.tune_clear()
    try:
        base64 = forces(mdgpu.FP64)
        rep = mdgpu.tune(ps, mdgpu.MIXED, context_ions=30000)
        rep64 = mdgpu.tune(ps, mdgpu.FP64, context_ions=30000)
        assert rep["classes"] == len(ps) and rep["runs"] >= len(ps)
        assert rep64["classes"] == len(ps)
        for p in ps:
            assert mdgpu.box_tiling(p, mdgpu.MIXED)["tuned"] == 1
            assert mdgpu.box_tiling(p, mdgpu.FP64)["tuned"] == 1
        t64, t32 = forces(mdgpu.FP64), forces(mdgpu.MIXED)
        for a, b, c in zip(t64, base64, t32):
            assert H.rel_err(a, b) <= 1e-12
            assert H.rel_err(c, a) <= MIXED_TOL
        # alone == in the batch under the tuned table
        b = mdgpu.Batch(ps, mdgpu.MIXED)
        for k, (r, v, sp) in enumerate(sts):
            b.set_state(k, r=r, v=v, species=sp)
        b.step(5)
        together = [b.state(k)["r"] for k in range(len(ps))]
        b.close()
        for k, p in enumerate(ps):
            lone = mdgpu.Batch([p], mdgpu.MIXED)
            r, v, sp = sts[k]
            lone.set_state(0, r=r, v=v, species=sp)
            lone.step(5)
            assert np.array_equal(lone.state(0)["r"], together[k])
            lone.close()
    finally:
        mdgpu.tune_clear()


def test_largest_box_kcl_32768_mixed_within_1e5(gpu):
    """The largest box of the job stream: KCl 16^3 cells (32768 ions,
    L = 100.6 A), rc = 10 A, eps 1e-6 (hm near 1958, hmax 44: three hz
    tiles of the structure factor, six K blocks of the back-projection,
    cell lists).  Mixed forces within 1e-5 of FP64 and the energy within
    1e-6; the FP64 path itself is pinned to the oracle at 4096 ions."""
    p = _params(16, H.A_KCL, salt="KCl")
    assert p.hm > 1500 and int(np.sqrt(p.hm)) >= 40
    r, v, sp = _state(p, 17, temp=1300.0, amp=0.2)
    res = {}
    for prec in (mdgpu.FP64, mdgpu.MIXED):
        b = mdgpu.Batch([p], prec)
        b.set_state(0, r=r, v=v, species=sp)
        b.forces()
        res[prec] = (b.state(0)["f"], b.scalars(0), b.work(0))
        b.close()
    (f64, s64, w64), (f32, s32, w32) = res[mdgpu.FP64], res[mdgpu.MIXED]
    assert w64["n_ions"] == 32768 and w32["pairs"] == w64["pairs"]
    assert H.rel_err(f32, f64) <= MIXED_TOL
    assert H.rms_rel_err(f32, f64) <= MIXED_TOL
    assert s32["etot"] == pytest.approx(s64["etot"], rel=1e-6)

