"""mdc's own program with its force path replaced at link time.

oracle/Makefile links one fixed mdc driver (oracle/dropin_main.c, mdc's
step loop of main.c:26-33) twice against the reference's sources:
mdc_cpu with every mdc routine, and mdc_gpu with real.c, rcprcl.c and
rv.c left out and libmdgpu.so linked in their place -- the recipe a
maintainer follows to adopt the GPU path.  Both print pot, kin, the
diagonal virial and the pressure every step, and position sums at the
end; the drop-in must agree with mdc to 1e-12 relative (the plain FP64
bar).  The programs are built from the reference checkout in this
container and travel prebuilt; without them the test is skipped.
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")
PROGS = {k: os.path.join(REF, k) for k in ("mdc_cpu", "mdc_gpu")}


def run(prog, n, steps, seed, loop=()):
    path = PROGS.get(prog, os.path.join(REF, prog))
    p = subprocess.run([path, str(n), str(steps), str(seed), *loop], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 0, (prog, p.returncode, p.stderr[-2000:])
    rows, msd, final, notes = [], [], None, 0
    for ln in p.stdout.splitlines():
        f = ln.split()
        if not f:
            continue
        if f[0] == "final":
            final = np.array([float(x) for x in f[1:]])
        elif f[0] == "msd":
            msd.append([float(x) for x in f[1:]])
        elif f[0].isdigit():
            rows.append([float(x) for x in f[1:]])
        else:   # control_temp's own report lines (pt.c:107-108)
            notes += 1
    return np.array(rows), final, p.stderr, np.array(msd), notes


@pytest.mark.skipif(not all(os.path.exists(p) for p in PROGS.values()),
                    reason="oracle/_ref drop-in programs not built (no reference checkout)")
@pytest.mark.parametrize("n,steps,loop", [(3, 20, ()), (4, 10, ()),
                                          (3, 30, ("5", "1500", "1", "10"))])
def test_mdc_program_linked_to_libmdgpu_matches_mdc(gpu, n, steps, loop):
    """loop: the rest of main.c's loop -- control_temp every 5 steps to
    1500 K, calc_msd every step, moment_correction every 10 steps."""
    cpu, fcpu, _, mcpu, ncpu = run("mdc_cpu", n, steps, 11, loop)
    dev, fdev, err, mdev, ndev = run("mdc_gpu", n, steps, 11, loop)
    if loop:
        assert ndev == ncpu == 2 * 6    # scalings at steps 5, 10, ..., 30
        assert mcpu.shape == mdev.shape == (steps, 3)
        assert np.array_equal(mdev[:, 0], mcpu[:, 0])
        # the MSD is a small difference of positions: the positions'
        # rounding-level differences show in it relatively magnified
        assert np.allclose(mdev[1:, 1:], mcpu[1:, 1:], rtol=1e-9, atol=0)
    assert "libmdgpu" not in err, err        # the drop-in reports failures on stderr
    assert cpu.shape == dev.shape == (steps, 6)
    # pot, kin, virX, virY, virZ, pres per step
    for c in range(6):
        scale = np.abs(cpu[:, c]).max()
        assert np.abs(dev[:, c] - cpu[:, c]).max() <= 1e-12 * scale, c
    assert np.allclose(fdev, fcpu, rtol=1e-12, atol=0)


# mdc's other tabulated ionic materials (mdc/examples/<m>: control.c,
# potential.h, real.c replacing program/'s), lattice constants from each
# example's get_control_param: the same drop-in, other tables, charges,
# lattices and ion counts (NaCl SX1 216, AgI 108, CaF2 324 ions)
MATERIALS = {"NaCl_sx1": 5.63, "AgI": 5.08, "CaF2": 5.9}


@pytest.mark.parametrize("mat", sorted(MATERIALS))
def test_other_tabulated_materials_match_mdc(gpu, mat):
    progs = [os.path.join(REF, f"mdc_{k}_{mat}") for k in ("cpu", "gpu")]
    if not all(os.path.exists(p) for p in progs):
        pytest.skip("oracle/_ref material programs not built (no reference checkout)")
    loop = ("0", "0", "0", "0", str(MATERIALS[mat]))
    cpu, fcpu, _, _, _ = run(f"mdc_cpu_{mat}", 3, 10, 5, loop)
    dev, fdev, err, _, _ = run(f"mdc_gpu_{mat}", 3, 10, 5, loop)
    assert "libmdgpu" not in err, err
    assert cpu.shape == dev.shape == (10, 6)
    for c in range(6):
        scale = np.abs(cpu[:, c]).max()
        assert np.abs(dev[:, c] - cpu[:, c]).max() <= 1e-12 * scale, (mat, c)
    assert np.allclose(fdev, fcpu, rtol=1e-12, atol=0)


@pytest.mark.skipif(not os.path.exists(PROGS["mdc_gpu"]),
                    reason="oracle/_ref drop-in programs not built (no reference checkout)")
def test_linked_program_exits_on_a_pair_inside_the_table_start(gpu):
    """A 0.9 A lattice puts nearest neighbours at about 0.45 A, inside the
    table's 0.5 A start, where mdc would index before its table
    (real.c:62).  The device skips such pairs and flags them; the drop-in
    then prints MDG_E_RANGE and exits with EXIT_FAILURE, mdc's convention
    for fatal errors, instead of letting mdc step on.  With
    MDGPU_DROPIN_ERRORS=status the program keeps running and the error is
    reported once."""
    args = [PROGS["mdc_gpu"], "3", "2", "11", "0", "0", "0", "0", "0.9"]
    env = dict(os.environ)
    env.pop("MDGPU_DROPIN_ERRORS", None)
    p = subprocess.run(args, capture_output=True, text=True, timeout=120, env=env)
    assert p.returncode == 1, (p.returncode, p.stderr[-2000:])
    assert "libmdgpu drop-in: next_rv_verlet failed" in p.stderr, p.stderr[-2000:]
    assert not any(ln.split()[:1] == ["1"] for ln in p.stdout.splitlines())
    env["MDGPU_DROPIN_ERRORS"] = "status"
    p = subprocess.run(args, capture_output=True, text=True, timeout=120, env=env)
    assert p.returncode == 0, (p.returncode, p.stderr[-2000:])
    assert p.stderr.count("libmdgpu drop-in") == 1, p.stderr[-2000:]
