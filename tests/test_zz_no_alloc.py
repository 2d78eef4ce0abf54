"""mdg_forces_noalloc, the allocation-free calc_foc, against mdg_batch_forces.

The entry point takes the caller's device buffers and keeps every other
device buffer in one static device array, so the test owns the device
memory: it allocates it through the HIP runtime libmdgpu itself links
(ctypes on libamdhip64), fills it, calls through the C ABI and compares
with a batch of the same box alone (mdg_batch_create + set_state +
mdg_batch_forces), which allocates.  Forces come from the same kernels on
the same work items, so they are bit-identical; the energies agree to
1e-13 (the self term is summed on the device in another order).

Two calls in a row, the second on a smaller box, then the first box again,
then a repeat (the cached plan), in both precisions; plus the refusals:
a species out of range, a non-finite coordinate and a box beyond the
scratch, each leaving the next call correct.
"""
import ctypes as C

import numpy as np
import pytest

import gpuhelp as H
import mdgpu
import mdref

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        mdgpu.lib()  # loads /opt/rocm's runtime first; the same one is reused
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


H2D, D2H = 1, 2


class Dev:
    """A device buffer holding a copy of a numpy array."""

    def __init__(self, a):
        self.a = np.ascontiguousarray(a)
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), max(self.a.nbytes, 8)) == 0
        assert hip().hipMemcpy(self.p, self.a.ctypes.data, self.a.nbytes, H2D) == 0

    def get(self):
        out = np.empty_like(self.a)
        assert hip().hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.p:
            hip().hipFree(self.p)
            self.p = C.c_void_p()


def params(n, prec):
    a = H.A_NACL
    L = n * a
    rc = min(10.0, L / 2)
    alpha, hm = mdgpu.ewald_params(rc, L, 1e-6)
    mode = mdgpu.TABLE if prec == mdgpu.FP64 else mdgpu.ANALYTIC
    return mdgpu.box_params(n, n, n, a, alpha, hm, rc, 2.0, real_mode=mode)


def state(n, seed):
    r, sp = mdgpu.rocksalt(n, n, n, H.A_NACL)
    return mdref.jitter_positions(r, np.full(3, n * H.A_NACL), 0.25, seed), sp.astype(np.int32)


def noalloc(p, prec, r, sp):
    n = sp.size
    dr, ds = Dev(r.reshape(3 * n)), Dev(sp)
    df, de = Dev(np.full(3 * n, np.nan)), Dev(np.full(8, np.nan))
    try:
        rc = mdgpu.lib().mdg_forces_noalloc(C.byref(p), prec, dr.p, ds.p, df.p, de.p, None)
        return rc, df.get().reshape(3, n), de.get()
    finally:
        for b in (dr, ds, df, de):
            b.free()


def batch(p, prec, r, sp):
    b = mdgpu.Batch([p], prec)
    try:
        b.set_state(0, r=r, species=sp)
        b.forces()
        s = b.scalars(0)
        e = np.array([s["pot"], s["pot_real"], s["pot_recip"], s["pot_self"], s["etot"],
                      *s["vir"]])
        return b.state(0)["f"], e
    finally:
        b.close()


def check_same(p, prec, r, sp):
    rc, f, e = noalloc(p, prec, r, sp)
    assert rc == mdgpu.OK, mdgpu.errstr(rc)
    fb, eb = batch(p, prec, r, sp)
    assert np.all(np.isfinite(f))
    assert np.array_equal(f, fb)
    assert np.allclose(e, eb, rtol=1e-13, atol=0), (e, eb)


@pytest.mark.parametrize("prec", [mdgpu.FP64, mdgpu.MIXED])
def test_noalloc_matches_batch_forces_on_calls_in_a_row(gpu, prec):
    big, small = (params(4, prec), state(4, 3)), (params(3, prec), state(3, 4))
    check_same(big[0], prec, *big[1])
    check_same(small[0], prec, *small[1])   # second call: a smaller box
    check_same(big[0], prec, *big[1])       # the first box again (plan rebuilt)
    r2, sp2 = state(4, 9)
    check_same(big[0], prec, r2, sp2)       # same box, new positions (plan reused)


def test_noalloc_fp64_matches_the_oracle(gpu):
    """The FP64 table path through the allocation-free call also meets the
    plain path's 1e-12 bar against the oracle (mdc's calc_foc)."""
    n = 3
    L = n * H.A_NACL
    o = mdref.Oracle(n, n, n, H.A_NACL, 5.6 / L, 27, L / 2, 2.0)
    r = mdref.jitter_positions(o.r, o.L, 0.25, 6)
    o.set_state(r=r, f=np.zeros_like(r))
    o.calc_foc()
    p = mdgpu.box_params(n, n, n, H.A_NACL, 5.6 / L, 27, L / 2, 2.0, real_mode=mdgpu.TABLE)
    rc, f, e = noalloc(p, mdgpu.FP64, r, o.ion.astype(np.int32))
    assert rc == mdgpu.OK
    assert H.rel_err(f, o.f) <= 1e-12
    assert e[0] == pytest.approx(o.b.pot, rel=1e-12)
    assert np.allclose(e[5:8], [o.b.virX, o.b.virY, o.b.virZ], rtol=1e-12, atol=0)


def test_noalloc_refusals_leave_the_next_call_correct(gpu):
    prec = mdgpu.MIXED
    p, (r, sp) = params(3, prec), state(3, 5)
    bad = sp.copy()
    bad[7] = 5
    rc, _, _ = noalloc(p, prec, r, bad)
    assert rc == mdgpu.E_ARG
    rn = r.copy()
    rn[1, 3] = np.nan
    rc, _, _ = noalloc(p, prec, rn, sp)
    assert rc == mdgpu.E_STATE
    # a box beyond the 32 MiB scratch is refused before anything runs
    p_big = params(12, prec)
    r_big, sp_big = state(12, 1)
    rc, _, _ = noalloc(p_big, prec, r_big, sp_big)
    assert rc == mdgpu.E_BUDGET
    check_same(p, prec, r, sp)
