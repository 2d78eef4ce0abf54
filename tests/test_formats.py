"""mdc's output formats (files.c) as libmdgpu writes them: the text of one
files/out line, one positions record and one velocities record must be
byte-identical to what mdc's own print_to_file, r_position and
r_velocities (compiled from the reference sources) write for the same
state.  Host-side formatting only: no device is used."""
import numpy as np
import pytest

import mdgpu
import mdref

A = 5.64

pytestmark = pytest.mark.skipif(not mdref.ref_available(),
                                reason="oracle/_ref not built (no reference checkout)")


@pytest.fixture(scope="module")
def ref_state():
    L = 3 * A
    R = mdref.Reference(3, 3, 3, A, 5.6 / L, 27, L / 2, 2.0)
    R.reset_control_temp()
    o = mdref.Oracle(3, 3, 3, A, 5.6 / L, 27, L / 2, 2.0)
    r = mdref.jitter_positions(o.r, o.L, 0.2, 91)
    v = mdref.maxwell_velocities(o.masses(), 1200.0, 91)
    R.set_state(r=r, v=v, f=np.zeros_like(r))
    R.newton(1, 12, t_control_step=5, temp=1500.0, moment_every=10, msd=True)
    return R, o


def test_out_line_matches_print_to_file(ref_state, tmp_path):
    R, o = ref_state
    want = R.write("print_to_file", tmp_path / "out")
    s = R.scalars()
    _, _, val, _ = R.msd()
    got = mdgpu.format_out(12, s["kin"], s["pot"], s["pres"], R.N, o.L, (3, 3, 3), msd=val)
    assert got == want
    assert got.startswith("12  ") and got.endswith("\n")


def test_positions_record_matches_r_position(ref_state, tmp_path):
    R, _ = ref_state
    want = R.write("r_position", tmp_path / "positions")
    got = mdgpu.format_positions(R.species(), R.state()["r"])
    assert got == want
    assert got.count(" \n") == R.N // 4


def test_velocities_record_matches_r_velocities(ref_state, tmp_path):
    R, _ = ref_state
    want = R.write("r_velocities", tmp_path / "velocities")
    got = mdgpu.format_velocities(R.species(), R.state()["v"])
    assert got == want


def test_ragged_last_line_and_empty_record():
    """N not a multiple of 4 ends without the four-ion break; N = 0 is a
    lone newline (r_position's loop does not run)."""
    sp = np.array([0, 1, 0, 1, 0], dtype=np.int32)
    r = np.arange(15, dtype=np.float64).reshape(3, 5) * 0.5
    t = mdgpu.format_positions(sp, r)
    assert t.count("\n") == 2 and t.endswith(" \n0 2.00000 4.50000 7.00000 \n")
    assert mdgpu.format_positions(np.zeros(0, np.int32), np.zeros((3, 0))) == "\n"
