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


def run(prog, n, steps, seed):
    p = subprocess.run([PROGS[prog], str(n), str(steps), str(seed)], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 0, (prog, p.returncode, p.stderr[-2000:])
    rows, final = [], None
    for ln in p.stdout.splitlines():
        f = ln.split()
        if f[0] == "final":
            final = np.array([float(x) for x in f[1:]])
        else:
            rows.append([float(x) for x in f[1:]])
    return np.array(rows), final, p.stderr


@pytest.mark.skipif(not all(os.path.exists(p) for p in PROGS.values()),
                    reason="oracle/_ref drop-in programs not built (no reference checkout)")
@pytest.mark.parametrize("n,steps", [(3, 20), (4, 10)])
def test_mdc_program_linked_to_libmdgpu_matches_mdc(gpu, n, steps):
    cpu, fcpu, _ = run("mdc_cpu", n, steps, 11)
    dev, fdev, err = run("mdc_gpu", n, steps, 11)
    assert "libmdgpu" not in err, err        # the drop-in reports failures on stderr
    assert cpu.shape == dev.shape == (steps, 6)
    # pot, kin, virX, virY, virZ, pres per step
    for c in range(6):
        scale = np.abs(cpu[:, c]).max()
        assert np.abs(dev[:, c] - cpu[:, c]).max() <= 1e-12 * scale, c
    assert np.allclose(fdev, fcpu, rtol=1e-12, atol=0)
