"""Read-back of device-resident boxes while they step
(mdg_batch_step_recorded) and mdc's output files written from it
(mdg_batch_write_mdc_files, the formats of files.c)."""
import os

import numpy as np
import pytest

import gpuhelp as H
import mdgpu
import mdref

pytestmark = pytest.mark.gpu


def _boxes():
    os_ = [H.cfg1_oracle(n=n) for n in (3, 2, 4)]
    ps = [mdgpu.box_params(o.b.nx, o.b.ny, o.b.nz, o.b.Ax, o.b.a1, o.b.hm, o.b.radius, 2.0)
          for o in os_]
    starts = []
    for k, o in enumerate(os_):
        r = mdref.jitter_positions(o.r, o.L, 0.1, 200 + k)
        v = mdref.maxwell_velocities(o.masses(), 1000.0, 200 + k)
        starts.append((r, v, o.ion))
    return ps, starts


def _batch(ps, starts, prec=mdgpu.FP64, arm=(0,)):
    b = mdgpu.Batch(ps, prec)
    for k, (r, v, _) in enumerate(starts):
        b.set_state(k, r=r, v=v)
    for k in arm:
        b.set_observe(k, msd=True)
    return b


@pytest.mark.parametrize("nsteps,every", [(12, 4), (32, 8)])
def test_frames_are_the_states_at_those_steps(gpu, nsteps, every):
    """Each frame holds exactly the state a twin batch reaches by stepping
    `every` at a time and reading back (eager steps for every = 4, graph
    replays for every = 8), and recording leaves the trajectory unchanged."""
    ps, starts = _boxes()
    rec = _batch(ps, starts, mdgpu.MIXED)
    frames = rec.step_recorded(nsteps, every)
    twin = _batch(ps, starts, mdgpu.MIXED)
    assert len(frames) == len(ps) * nsteps // every
    it = iter(frames)
    for q in range(nsteps // every):
        twin.step(every)
        for k in range(len(ps)):
            f = next(it)
            assert f["box"] == k and f["step"] == (q + 1) * every
            st = twin.state(k)
            assert np.array_equal(f["r"], st["r"]) and np.array_equal(f["v"], st["v"])
            assert f["scalars"] == twin.scalars(k)
            assert np.array_equal(f["msd"], twin.observed(k)["msd"])
    assert frames[-3]["msd"][0] > 0 and not frames[-2]["msd"].any()   # box 0 armed, box 1 not
    for k in range(len(ps)):
        assert np.array_equal(rec.state(k)["r"], twin.state(k)["r"])
    rec.close()
    twin.close()


def test_copy_only_recording_keeps_the_trajectory(gpu):
    ps, starts = _boxes()
    a = _batch(ps, starts)
    b = _batch(ps, starts)
    assert a.step_recorded(24, 6, keep=False) == []
    b.step(24)
    for k in range(len(ps)):
        for key in ("r", "v", "f"):
            assert np.array_equal(a.state(k)[key], b.state(k)[key])
    a.close()
    b.close()


def test_mdc_files_hold_mdc_records_of_the_frames(gpu, tmp_path):
    """files/out, positions and velocities per box: mdc's blank, then one
    record per frame in files.c's format (pinned byte for byte against the
    compiled files.c in tests/test_formats.py) of the state at that step."""
    ps, starts = _boxes()
    for k in range(len(ps)):
        os.makedirs(tmp_path / f"box{k}")
    w = _batch(ps, starts)
    w.write_mdc_files(10, 5, mdgpu.WRITE_POSITIONS | mdgpu.WRITE_VELOCITIES, str(tmp_path))
    twin = _batch(ps, starts)
    want = {k: {"out": " ", "positions": " ", "velocities": " "} for k in range(len(ps))}
    for q in range(2):
        twin.step(5)
        for k, p in enumerate(ps):
            st, sc = twin.state(k), twin.scalars(k)
            sp = starts[k][2]
            L = (p.nx * p.ax, p.ny * p.ay, p.nz * p.az)
            want[k]["out"] += mdgpu.format_out(sc["step"], sc["kin"], sc["pot"], sc["pres"],
                                               sp.size, L, (p.nx, p.ny, p.nz),
                                               msd=twin.observed(k)["msd"][:2])
            want[k]["positions"] += mdgpu.format_positions(sp, st["r"])
            want[k]["velocities"] += mdgpu.format_velocities(sp, st["v"])
    for k in range(len(ps)):
        for name, text in want[k].items():
            assert open(tmp_path / f"box{k}" / name).read() == text, (k, name)
    assert open(tmp_path / "box0" / "out").read().splitlines()[1].startswith("10  ")
    w.close()
    twin.close()
