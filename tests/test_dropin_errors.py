"""The drop-in's error convention, on the CPU.

mdc exits on a fatal error (mdc/program/files.c:9-24) and its main.c
checks no status, so a failing drop-in routine prints the error and exits
with EXIT_FAILURE instead of letting mdc step on with arrays it never
updated.  A host that checks mdg_dropin_status opts out with
MDGPU_DROPIN_ERRORS=status.  Without a GPU every routine fails at its first
call (no device), which exercises both modes through mdc's own program
linked to libmdgpu (oracle/_ref/mdc_gpu, built from the reference
checkout)."""
import os
import subprocess

import pytest

import mdgpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "oracle", "_ref", "mdc_gpu")

needs_prog = pytest.mark.skipif(not os.path.exists(PROG),
                                reason="oracle/_ref/mdc_gpu not built (no reference checkout)")


def _no_device():
    try:
        return mdgpu.device_count() == 0
    except Exception:
        return True


@needs_prog
@pytest.mark.skipif(not _no_device(), reason="a GPU is visible: the routines succeed")
def test_linked_mdc_program_exits_on_the_first_failure():
    env = dict(os.environ)
    env.pop("MDGPU_DROPIN_ERRORS", None)
    p = subprocess.run([PROG, "3", "5", "1"], capture_output=True, text=True, timeout=60,
                       env=env)
    assert p.returncode == 1, (p.returncode, p.stderr[-500:])
    assert "libmdgpu drop-in" in p.stderr and "failed" in p.stderr
    # it stopped at the first routine: no step line was printed
    assert not any(ln.split()[:1] == ["1"] for ln in p.stdout.splitlines())


@needs_prog
@pytest.mark.skipif(not _no_device(), reason="a GPU is visible: the routines succeed")
def test_status_mode_keeps_the_program_running():
    env = dict(os.environ, MDGPU_DROPIN_ERRORS="status")
    p = subprocess.run([PROG, "3", "5", "1"], capture_output=True, text=True, timeout=60,
                       env=env)
    assert p.returncode == 0, (p.returncode, p.stderr[-500:])
    assert p.stderr.count("libmdgpu drop-in") == 1   # reported once
    assert sum(ln.split()[:1] == ["5"] for ln in p.stdout.splitlines()) == 1


def test_error_mode_setter_rejects_unknown_modes():
    L = mdgpu.lib()
    assert L.mdg_dropin_set_error_mode(7) == mdgpu.E_ARG
    assert L.mdg_dropin_set_error_mode(mdgpu.DROPIN_STATUS) == mdgpu.OK
