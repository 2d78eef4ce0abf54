"""libmdgpu_pr.so -- pr-mdc's force routines on the GPU -- against the
oracle's restatement of pr-mdc (oracle/pr_oracle.c, pinned bit for bit to
pr-mdc's compiled sources) and against pr-mdc's own program.

Every routine is called through the C ABI over pr-mdc's globals (bound to
numpy arrays), on NaCl systems in cubic and sheared (triclinic) cells of
64 to 1728 ions.  The bar is the plain FP64 one: forces, energy and the
virial tensor within 1e-12 of the oracle relative to their largest
magnitude; with one split the table routine's forces are bit-identical."""
import os
import subprocess

import numpy as np
import pytest

import prmdgpu
import prref

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")

SYSTEMS = [(2, (0.0, 0.0, 0.0), 1), (2, (0.1, 0.05, -0.08), 2), (3, (-0.2, 0.15, 0.1), 3),
           (4, (0.15, -0.1, 0.2), 4), (6, (0.05, 0.1, -0.12), 5)]


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


def gpu_run(s, routines):
    """Binds s into libmdgpu_pr's globals, runs the routines, returns
    (forces, pot, virial)."""
    L = prmdgpu.lib()
    g = s.f.copy()
    keep = s.f
    s.f = g
    s.virial.a[:] = 0.0
    pot0 = s.pot
    s.bind(L)
    try:
        for n in routines:
            print(f"prmdc {n} N={s.N}", flush=True)
            prmdgpu.call(n)
        pot = s.read(L)
        vir = s.virial.a.copy()
    finally:
        s.f, s.pot = keep, pot0
        s.virial.a[:] = 0.0
    return g, pot, vir


@pytest.fixture(autouse=True)
def auto_splits():
    prmdgpu.set_splits(0, 0)
    yield
    prmdgpu.set_splits(0, 0)


@pytest.mark.parametrize("n,shear,seed", SYSTEMS)
def test_each_routine_matches_oracle(gpu, n, shear, seed):
    s = prref.pr_system(n=n, shear=shear, seed=seed)
    done = []
    for name in prref.ROUTINES:
        done.append(name)
        f, pot, vir = gpu_run(s, done)
        f0, pot0, vir0 = prref.oracle_run(s, done)
        assert rel(f, f0) < 1e-12, (name, rel(f, f0))
        assert abs(pot - pot0) <= 1e-12 * abs(pot0), name
        assert rel(vir, vir0) < 1e-12, name
        assert np.array_equal(vir, vir.T), name


@pytest.mark.parametrize("n,shear,seed", SYSTEMS[1:4])
def test_one_split_table_forces_are_bit_identical(gpu, n, shear, seed):
    s = prref.pr_system(n=n, shear=shear, seed=seed)
    prmdgpu.set_splits(1, 1)
    f, pot, vir = gpu_run(s, ("real_space",))
    f0, pot0, vir0 = prref.oracle_run(s, ("real_space",))
    assert np.array_equal(f, f0)
    assert abs(pot - pot0) <= 1e-13 * abs(pot0)
    f, pot, vir = gpu_run(s, prref.ROUTINES)
    f1, _, _ = prref.oracle_run(s, prref.ROUTINES)
    assert rel(f, f1) < 1e-12


def test_splits_do_not_change_results(gpu):
    s = prref.pr_system(n=4, shear=(0.1, 0.05, -0.08), seed=8)
    ref = gpu_run(s, prref.ROUTINES)
    for pair, ks in ((1, 1), (3, 7), (64, 309)):
        prmdgpu.set_splits(pair, ks)
        out = gpu_run(s, prref.ROUTINES)
        assert rel(out[0], ref[0]) < 1e-13, (pair, ks)
        assert abs(out[1] - ref[1]) <= 1e-13 * abs(ref[1])


def test_reciprocal_cutoffs_and_alpha(gpu):
    """Other hm and alpha than pr-mdc's defaults, and the skip of hm = 0."""
    s = prref.pr_system(n=3, shear=(0.1, -0.05, 0.08), seed=6)
    for a1, hmax in ((0.3, 50), (0.9, 100), (0.5, 0)):
        s.a1, s.hmax = a1, hmax
        f, pot, vir = gpu_run(s, ("ewald1", "ewald2", "ewald3"))
        f0, pot0, vir0 = prref.oracle_run(s, ("ewald1", "ewald2", "ewald3"))
        assert rel(f, f0) < 1e-12 and abs(pot - pot0) <= 1e-12 * abs(pot0), (a1, hmax)


def test_empty_system(gpu):
    s = prref.pr_system(n=2)
    s.N = 0
    f, pot, vir = gpu_run(s, prref.ROUTINES)
    assert pot == 0.0 and not vir.any()


def test_pair_inside_table_start_is_an_error(gpu):
    s = prref.pr_system(n=2, seed=3)
    s.r[:, 1] = s.r[:, 0] + np.array([0.3, 0.0, 0.0])   # 0.3 A apart: before the table's 0.5 A
    with pytest.raises(prmdgpu.MdgPrError) as e:
        gpu_run(s, ("real_space",))
    assert e.value.code == prmdgpu.E_RANGE
    assert prmdgpu.lib().mdg_pr_status() == 0          # cleared by the raise
    gpu_run(prref.pr_system(n=2), ("real_space",))      # and the library still works


def test_profile_times_resident_evaluation(gpu):
    s = prref.pr_system(n=4, shear=(0.1, 0.05, -0.08), seed=8)
    gpu_run(s, prref.ROUTINES)
    ms, per = prmdgpu.profile(5)
    assert ms > 0 and len(per) == 4 and all(p > 0 for p in per)
    assert ms == pytest.approx(sum(per), rel=1e-9)


def run(prog, *args):
    p = subprocess.run([os.path.join(REF, prog), *map(str, args)], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, (prog, p.stderr[-2000:])
    rows = [ln.split() for ln in p.stdout.splitlines() if ln.strip()]
    steps = np.array([[float(x) for x in r[1:]] for r in rows if r[0].isdigit()])
    final = np.array([float(x) for x in rows[-1][1:]]) if rows[-1][0] == "final" else None
    return steps, final, p.stderr


@pytest.mark.skipif(not all(os.path.exists(os.path.join(REF, p)) for p in ("prmdc_cpu",
                                                                           "prmdc_gpu")),
                    reason="oracle/_ref pr-mdc programs not built (no reference checkout)")
@pytest.mark.parametrize("n,steps,shear", [(2, 20, (0.1, 0.05, -0.08)), (3, 20, (0.0, 0.0, 0.0)),
                                           (3, 15, (-0.2, 0.15, 0.1))])
def test_prmdc_program_linked_to_libmdgpu_pr_matches_prmdc(gpu, n, steps, shear):
    """pr-mdc's NPT step loop (Parrinello-Rahman cell dynamics, triclinic
    cell) with real.c and ewald.c replaced by libmdgpu_pr."""
    print(f"prmdc program n={n} steps={steps} shear={shear}", flush=True)
    a, fa, _ = run("prmdc_cpu", n, steps, 17, *shear)
    b, fb, err = run("prmdc_gpu", n, steps, 17, *shear)
    assert "libmdgpu_pr" not in err, err
    assert a.shape == b.shape == (steps, 13)
    for c in range(a.shape[1]):
        assert rel(b[:, c], a[:, c]) < 1e-10, c
    assert rel(fb, fa) < 1e-10
    # the cell did move (constant pressure), and it stayed triclinic if it started so
    assert not np.array_equal(a[0, 4:], a[-1, 4:])


@pytest.mark.skipif(not (os.path.exists(prref.PREF_SO) and
                         os.path.exists(os.path.join(REF, "prmdc_cpu"))),
                    reason="oracle/_ref pr-mdc harness and program not built")
@pytest.mark.parametrize("n,steps,shear", [(2, 25, (0.1, 0.05, -0.08)), (3, 20, (-0.2, 0.15, 0.1)),
                                           (3, 12, (0.0, 0.0, 0.0)), (2, 45, (0.1, 0.05, -0.08))])
def test_resident_npt_run_matches_prmdc_program(gpu, n, steps, shear):
    """mdg_pr_run: pr-mdc's whole step loop (forces, velocity Verlet with the
    scaled-coordinate wrap, Parrinello-Rahman cell dynamics, the thermostat
    at step 10 and 20, kinetic and pressure tensors) resident on the
    device, against pr-mdc's own program from the same start."""
    print(f"prmdc resident run n={n} steps={steps} shear={shear}", flush=True)
    R = prref.ref()
    R.prref_setup(n, 5.63, *shear, 17, 0.2, 0.05)
    L = prmdgpu.lib()
    prmdgpu.copy_globals(R, L)
    rec = prmdgpu.run(steps)
    a, fa, _ = run("prmdc_cpu", n, steps, 17, *shear)
    assert rec.shape == a.shape
    for c in range(a.shape[1]):
        assert rel(rec[:, c], a[:, c]) < (1e-10 if steps <= 25 else 1e-9), c
    ctl, cell, sys_, pt, ion = prmdgpu.globals_of(L)
    N = sys_.N
    sums = [np.ctypeslib.as_array(p, shape=(N,)).sum() for p in
            (cell.rx, cell.ry, cell.rz, cell.vx, cell.vy, cell.vz)]
    assert rel(np.array(sums), fa) < 1e-9
    assert sys_.step == steps


@pytest.mark.skipif(not os.path.exists(prref.PREF_SO), reason="pr-mdc harness not built")
def test_resident_graph_blocks_equal_single_steps(gpu):
    """A 60-step run replays a 16-step hipGraph three times between eager
    steps; 60 one-step calls launch every kernel eagerly, with the state
    round-tripping through pr-mdc's globals in between.  Thermostat off (its
    running average restarts per call): the two are bit-identical."""
    R = prref.ref()
    L = prmdgpu.lib()
    out = []
    for mode in ("graph", "eager"):
        R.prref_setup(2, 5.63, 0.1, 0.05, -0.08, 23, 0.2, 0.05)
        prmdgpu.globals_of(R)[3].t_control_step = 0
        prmdgpu.copy_globals(R, L)
        if mode == "graph":
            rec = prmdgpu.run(60)
        else:
            rec = np.concatenate([prmdgpu.run(1) for _ in range(60)])
        ctl, cell, sys_, pt, ion = prmdgpu.globals_of(L)
        r = np.stack([np.ctypeslib.as_array(p, shape=(sys_.N,)).copy()
                      for p in (cell.rx, cell.ry, cell.rz, cell.vx, cell.vy, cell.vz)])
        out.append((rec, r))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_cutoff_past_table_end_is_an_error(gpu):
    """pr-mdc's cutoff is half the smallest face distance; past 10 unit
    cells it passes the table's 20 A end, where pr-mdc reads past its
    table.  The library refuses instead."""
    s = prref.pr_system(n=8, jitter=0.1)
    assert s.cutoff > 20.0
    with pytest.raises(prmdgpu.MdgPrError) as e:
        gpu_run(s, ("real_space",))
    assert e.value.code == prmdgpu.E_RANGE
    gpu_run(prref.pr_system(n=2), ("real_space",))


def test_run_without_cell_dynamics_state_is_an_argument_error(gpu):
    s = prref.pr_system(n=2)
    L = prmdgpu.lib()
    s.bind(L)
    pr = prmdgpu.PrSet.in_dll(L, "pr")
    keep = (pr.mass, pr.v)
    pr.mass, pr.v = 0.0, None
    try:
        with pytest.raises(prmdgpu.MdgPrError) as e:
            prmdgpu.run(5)
        assert e.value.code == prmdgpu.E_ARG
    finally:
        pr.mass, pr.v = keep
