"""Pins the pr-mdc part of the oracle (oracle/pr_oracle.c): against pr-mdc's
own routines compiled from its sources (bit for bit), against closed forms
(k-vector counts, the NaCl Madelung constant in a cubic and in a sheared
cell of the same crystal) and against invariants (forces are minus the
gradient of the energy, the virial tensor is minus its strain derivative).
CPU only."""
import ctypes as C
import os

import numpy as np
import pytest

import prref

MADELUNG_NACL = 1.747564594633182  # rocksalt, referred to the nearest-neighbour distance


needs_ref = pytest.mark.skipif(not os.path.exists(prref.PREF_SO),
                               reason="pr-mdc reference harness not built")


def ref_state(R):
    ctl, cell, sys_, pt, ion = prref.prmdgpu.globals_of(R)
    N = sys_.N
    f = np.stack([np.ctypeslib.as_array(p, shape=(N,)).copy() for p in (cell.fx, cell.fy, cell.fz)])
    vir = np.array([[pt.virial[i][j] for j in range(3)] for i in range(3)])
    return f, sys_.pot, vir


@needs_ref
@pytest.mark.parametrize("n,shear,seed", [(2, (0.0, 0.0, 0.0), 3), (2, (0.1, 0.05, -0.08), 7),
                                          (3, (-0.2, 0.15, 0.1), 11)])
def test_oracle_is_compiled_prmdc_bit_for_bit(n, shear, seed):
    R = prref.ref()
    N = R.prref_setup(n, 5.63, *shear, seed, 0.2, 0.05)
    assert N == 8 * n ** 3
    s = prref.system_from_ref(R)
    R.clear_foc()
    s.f[:] = 0.0
    # the oracle's cell algebra from H alone reproduces pr-mdc's
    b = prref.box_of(s)
    b.hmr[:] = [0.0] * 9
    assert prref.orc().orc_pr_calc_cell(C.byref(b)) == 0
    prref.orc().orc_pr_calc_sigma(C.byref(b))
    assert np.array_equal(np.array(b.hmr[:]), s.hmr.a.ravel())
    assert np.array_equal(np.array(b.g[:]), s.g.a.ravel()) and b.vol == s.vol
    prref.orc().orc_pr_set_cutoff(C.byref(b), 0, 0)
    assert (b.cutoff, b.a1) == (s.cutoff, s.a1)
    done = []
    for name in prref.ROUTINES:
        getattr(R, name)()
        done.append(name)
        f_ref, pot_ref, vir_ref = ref_state(R)
        f, pot, vir = prref.oracle_run(s, done)
        assert np.array_equal(f, f_ref), name
        assert pot == pot_ref, name
        assert np.array_equal(vir, vir_ref), name


@needs_ref
def test_min_image_is_cyclic_bc_dr_bit_for_bit():
    R = prref.ref()
    R.prref_setup(2, 5.63, 0.1, 0.05, -0.08, 5, 0.2, 0.05)
    s = prref.system_from_ref(R)
    b = prref.box_of(s)
    rng = np.random.default_rng(2)
    H = s.hm.a
    vecs = list(rng.uniform(-12, 12, size=(200, 3)))
    # scaled components exactly on the fold boundaries
    for sc in ([0.5, 0.0, 0.0], [-0.5, 0.25, 0.0], [0.5, -0.5, 0.5], [0.49999, 0.5, -0.50001]):
        vecs.append(H.T @ np.array(sc))
    for v in vecs:
        a = [C.c_double(x) for x in v]
        o = [C.c_double(x) for x in v]
        R.cyclic_bc_dr(*[C.byref(x) for x in a])
        prref.orc().orc_pr_min_image(C.byref(b), *[C.byref(x) for x in o])
        assert [x.value for x in a] == [x.value for x in o]


def test_kvector_counts():
    # manual.IEMD:250-253: hm = 20 gives 194 half-sphere vectors; hm = 27, 309
    assert prref.orc().orc_pr_count_kvectors(20, 4) == 194
    assert prref.orc().orc_pr_count_kvectors(27, 5) == 309


def coulomb(s, a1, hmax):
    """E1 + E2/4 + E3 at alpha a1, the textbook Ewald energy (pr-mdc's
    ewald2 energy is four times the textbook reciprocal term; its force is
    the textbook one)."""
    s.a1, s.hmax = a1, hmax
    _, e1, _ = prref.oracle_run(s, ("ewald1",))
    _, e2, _ = prref.oracle_run(s, ("ewald2",))
    _, e3, _ = prref.oracle_run(s, ("ewald3",))
    return e1 + e2 / 4.0 + e3, (e1, e2, e3)


def test_madelung_constant_cubic():
    s = prref.pr_system(n=2, jitter=0.0)
    a = 5.63
    a1 = 6.0 / s.cutoff                      # erfc(6) ~ 2e-17 at the cutoff
    e, (e1, e2, e3) = coulomb(s, a1, 24 * 24)
    want = -(s.N / 2) * MADELUNG_NACL * prref.KK / (a / 2)
    assert e == pytest.approx(want, rel=1e-11)
    # the energy as pr-mdc sums it carries the reciprocal term four times
    assert abs((e1 + e2 + e3) - want) > 1e-3 * abs(want)


def test_madelung_constant_sheared_cell():
    """The same crystal in a triclinic supercell (b' = b + a, c' = c + b):
    a transposed g, a wrong fold or a wrong volume changes the sum."""
    s = prref.pr_system(n=2, jitter=0.0, shear=(1.0, 0.0, 1.0), same_crystal=True)
    a = 5.63
    assert s.cutoff < 3.3                   # half the smallest face distance L/sqrt(3)
    a1 = 4.8 / s.cutoff
    e, _ = coulomb(s, a1, 48 * 48)
    want = -(s.N / 2) * MADELUNG_NACL * prref.KK / (a / 2)
    assert e == pytest.approx(want, rel=2e-9)


def test_sheared_cell_of_orthorhombic_crystal():
    """A stretched (orthorhombic) rocksalt in its own cell and in a sheared
    cell: equal energies.  Unlike the cubic crystal, whose reciprocal
    lattice is the same set either way, this fails for a transposed g."""
    st = (1.0, 1.13, 0.91)
    o = prref.pr_system(n=2, jitter=0.0, stretch=st)
    e_o, _ = coulomb(o, 5.0 / o.cutoff, 30 * 30)
    s = prref.pr_system(n=2, jitter=0.0, stretch=st, shear=(1.0, 0.0, 1.0), same_crystal=True)
    e_s, _ = coulomb(s, 4.8 / s.cutoff, 48 * 48)
    assert e_s == pytest.approx(e_o, rel=2e-9)


def total_energy(s, r):
    keep = s.r.copy()
    s.r[:] = r
    _, e1, _ = prref.oracle_run(s, ("ewald1",))
    _, e2, _ = prref.oracle_run(s, ("ewald2",))
    _, e3, _ = prref.oracle_run(s, ("ewald3",))
    s.r[:] = keep
    return e1 + e2 / 4.0 + e3


def test_ewald_forces_are_minus_gradient_triclinic():
    s = prref.pr_system(n=2, shear=(0.1, 0.05, -0.08), seed=4)
    s.a1 = 5.5 / s.cutoff
    f, _, _ = prref.oracle_run(s, ("ewald1", "ewald2", "ewald3"))
    h = 1e-5
    for i in (0, 17, 40, 63):
        for c in range(3):
            rp, rm = s.r.copy(), s.r.copy()
            rp[c, i] += h
            rm[c, i] -= h
            g = (total_energy(s, rp) - total_energy(s, rm)) / (2 * h)
            assert -g == pytest.approx(f[c, i], rel=1e-6, abs=1e-7 * np.abs(f).max())


def test_ewald1_virial_is_strain_derivative():
    """W_ab = sum_pairs f_ij d_a d_b = -dE1/de_ab for a symmetric strain e
    applied to the cell and the positions alike."""
    s = prref.pr_system(n=2, shear=(0.1, 0.05, -0.08), seed=9)
    s.a1 = 5.5 / s.cutoff
    _, _, W = prref.oracle_run(s, ("ewald1",))
    H0, r0 = s.hm.a.copy(), s.r.copy()

    def energy(eps):
        D = np.eye(3) + eps
        s.hm.a[:] = H0 @ D.T
        s.hmr.a[:] = np.linalg.inv(s.hm.a)
        s.r[:] = D @ r0
        _, e, _ = prref.oracle_run(s, ("ewald1",))
        s.hm.a[:] = H0
        s.hmr.a[:] = np.linalg.inv(H0)
        s.r[:] = r0
        return e

    h = 1e-6
    for a, b in ((0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)):
        E = np.zeros((3, 3))
        E[a, b] = E[b, a] = h
        d = (energy(E) - energy(-E)) / (2 * h)
        want = -(W[a, b] + W[b, a]) if a != b else -W[a, a]
        assert d == pytest.approx(want, rel=1e-6, abs=1e-6 * np.abs(W).max())
