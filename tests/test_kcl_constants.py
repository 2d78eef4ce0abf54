"""Pins for the KCl constants, which mdc does not hold.

mdc ships only NaCl's Tosi-Fumi set (mdc/program/control.c:158-180) and
cites the Tosi-Fumi papers for it (mdc/program/note:10-14).  The KCl set of
mdg_potential_kcl / orc_potential_kcl is taken from the same papers (Fumi &
Tosi, J. Phys. Chem. Solids 25, 31 (1964); Tosi & Fumi, ibid. 45) and put
into mdc's Born-Mayer-Huggins form.  Nothing in the reference can check
it, so it is pinned by what the physics fixes, with NaCl -- the reference's
own constants -- as the control throughout:

* closed forms: the Pauling factors c_ij = 1 + z_i/n_i + z_j/n_j of two
  closed 8-electron shells; sigma_ij = sigma_i + sigma_j with the chloride
  radius and the prefactor b shared with mdc's NaCl set; the unit
  conversion of C and D from 1e-60 erg cm^6 and 1e-76 erg cm^8, the one
  that turns the papers' NaCl values into control.c:175-180;
* the crystal: a smooth lattice sum (Madelung energy plus the BMH terms
  over lattice shells) of the static rock-salt energy; its minimum, energy
  and curvature are compared with the measured lattice constant, the
  Born-Haber lattice energy and the bulk modulus of the 4 K elastic
  constants;
* each plausible mistake (C and D not converted, NaCl's rho or sigma_Na
  carried over, b off by ten, a cross radius that is not a sum) is planted
  in a copy of the constants and must fail one of these checks;
* the oracle's own crystal energy (Ewald + analytic pairs) for both salts
  equals the lattice sum to 1e-9.
"""
import ctypes as C

import numpy as np
import pytest
from scipy.optimize import minimize_scalar

import mdgpu
import mdref

KK = 2.307079556            # e^2 / (4 pi eps0) in mdc's units (md.h)
MADELUNG = 1.7475645946332  # rock salt, per ion pair, nearest-neighbour distance
KJ_PER_MOL = 602.2136736    # 1e-18 J per ion pair -> kJ / mol

# measured crystal data (room-temperature lattice constant [A], Born-Haber
# lattice energy [kJ/mol], bulk modulus (c11 + 2 c12) / 3 of the 4.2 K
# elastic constants [GPa]: NaCl 57.33 / 11.23 (Lewis, Lehoczky & Briscoe,
# Phys. Rev. 161, 877 (1967)), KCl 48.33 / 5.42 (Norwood & Briscoe,
# Phys. Rev. 112, 45 (1958)))
CRYSTAL = {"NaCl": (5.6402, 787.0, (57.33 + 2 * 11.23) / 3),
           "KCl": (6.2917, 717.0, (48.33 + 2 * 5.42) / 3)}

_g = np.arange(-14, 15)
_l, _m, _n = np.meshgrid(_g, _g, _g, indexing="ij")
_D = np.sqrt(_l * _l + _m * _m + _n * _n).ravel()
_P = ((_l + _m + _n) % 2).ravel()
_D, _P = _D[_D > 0], _P[_D > 0]


class Pot:
    """A plain copy of mdg_potential's fields (mutable for planted mistakes)."""

    def __init__(self, p):
        self.Z = np.array(list(p.Z))
        self.m = np.array(list(p.m))
        self.b, self.rho = p.b, p.rho
        self.c = np.array(list(p.c))
        self.sigma = np.array(list(p.sigma))
        self.C = np.array(list(p.C))
        self.D = np.array(list(p.D))


def lattice_energy(pp, a, rmax=40.0):
    """Static rock-salt energy per ion pair [1e-18 J]: Madelung energy plus
    the BMH pair terms of mdc's form (control.c:219-240) summed over the
    lattice shells within rmax of an ion of each species."""
    r0 = a / 2.0
    R = _D * r0
    sel = R <= rmax
    R, par = R[sel], _P[sel]

    def phi(t):
        return (pp.c[t] * pp.b * np.exp((pp.sigma[t] - R) / pp.rho) - pp.C[t] / R ** 6
                - pp.D[t] / R ** 8)

    same = par == 0
    short = 0.5 * (np.sum((phi(0) + phi(2))[same]) + 2.0 * np.sum(phi(1)[~same]))
    return -MADELUNG * pp.Z[0] * -pp.Z[1] / r0 + short


def crystal(pp):
    """(a0 [A], E0 [kJ/mol], B0 [GPa]) at the static energy minimum."""
    res = minimize_scalar(lambda a: lattice_energy(pp, a), bounds=(4.0, 9.0),
                          method="bounded", options={"xatol": 1e-10})
    a, h = res.x, 1e-3
    d2 = (lattice_energy(pp, a + h) - 2 * lattice_energy(pp, a) + lattice_energy(pp, a - h)) / h ** 2
    # V = a^3 / 4 per ion pair: B = V d2E/dV2 = 4 E''(a) / (9 a); 1e-18 J / A^3 = 1e3 GPa
    return a, lattice_energy(pp, a) * KJ_PER_MOL, 4.0 * d2 / (9.0 * a) * 1e3


def crystal_ok(salt, pp):
    a_rt, e_bh, b_0k = CRYSTAL[salt]
    a0, e0, b0 = crystal(pp)
    # 0 -> 300 K thermal expansion of an alkali halide is about 1%
    return (0.985 * a_rt <= a0 <= a_rt and abs(-e0 - e_bh) <= 0.03 * e_bh
            and abs(b0 - b_0k) <= 0.04 * b_0k)


def pauling(z_i, n_i, z_j, n_j):
    return 1.0 + z_i / n_i + z_j / n_j


def oracle_potential(salt):
    p = mdref.OrcPotential()
    fn = {"NaCl": mdref.lib().orc_potential_nacl, "KCl": mdref.lib().orc_potential_kcl}[salt]
    fn(KK, C.byref(p))
    return p


@pytest.mark.parametrize("salt", ["NaCl", "KCl"])
def test_library_and_oracle_hold_the_same_constants(salt):
    lib, orc = Pot(mdgpu.potential(salt)), Pot(oracle_potential(salt))
    for f in ("Z", "m", "c", "sigma", "C", "D"):
        assert np.array_equal(getattr(lib, f), getattr(orc, f)), f
    assert lib.b == orc.b and lib.rho == orc.rho


def test_closed_forms_fix_the_kcl_set():
    na, k = Pot(mdgpu.potential("NaCl")), Pot(mdgpu.potential("KCl"))
    # Pauling factors: K+ and Cl- (like Na+) have closed 8-electron shells;
    # the formula reproduces mdc's NaCl factors (control.c:169-171)
    for pp in (na, k):
        assert np.allclose(pp.c, [pauling(1, 8, 1, 8), pauling(1, 8, -1, 8),
                                  pauling(-1, 8, -1, 8)], rtol=0, atol=1e-15)
    # sigma_ij = sigma_i + sigma_j, with Cl-'s radius as in mdc's NaCl set
    for pp in (na, k):
        s_c, s_a = pp.sigma[0] / 2, pp.sigma[2] / 2
        assert pp.sigma[1] == pytest.approx(s_c + s_a, rel=1e-15)
    assert k.sigma[2] == na.sigma[2]
    # the papers use one b for every alkali halide; Cl-'s mass is mdc's
    assert k.b == na.b and k.m[1] == na.m[1]
    assert k.Z[0] == na.Z[0] and k.Z[1] == na.Z[1]
    # C, D: 1e-60 erg cm^6 = 1e-67 J * 1e48 A^6 = 0.1 x (1e-18 J A^6), and
    # 1e-76 erg cm^8 likewise; the factor turns the papers' NaCl values into
    # control.c:175-180, and their KCl values into the library's
    f6 = 1e-60 * 1e-7 * (1e8) ** 6 / 1e-18
    f8 = 1e-76 * 1e-7 * (1e8) ** 8 / 1e-18
    assert np.allclose(na.C, np.array([1.68, 11.2, 116.0]) * f6, rtol=1e-12, atol=0)
    assert np.allclose(na.D, np.array([0.8, 13.9, 233.0]) * f8, rtol=1e-12, atol=0)
    assert np.allclose(k.C, np.array([24.3, 48.0, 124.5]) * f6, rtol=1e-12, atol=0)
    assert np.allclose(k.D, np.array([24.0, 73.0, 250.0]) * f8, rtol=1e-12, atol=0)
    # K's standard atomic weight 39.0983 g/mol in mdc's mass unit (control.c:161)
    assert k.m[0] * (6.0221367 * 1.6605402 / 10.0) == pytest.approx(39.0983, rel=1e-14)


@pytest.mark.parametrize("salt", ["NaCl", "KCl"])
def test_static_crystal_matches_measured_data(salt):
    a0, e0, b0 = crystal(Pot(mdgpu.potential(salt)))
    a_rt, e_bh, b_0k = CRYSTAL[salt]
    assert 0.985 * a_rt <= a0 <= a_rt, a0
    assert abs(-e0 - e_bh) <= 0.03 * e_bh, e0
    assert abs(b0 - b_0k) <= 0.04 * b_0k, b0
    assert crystal_ok(salt, Pot(mdgpu.potential(salt)))


def _mutations():
    def cd_unconverted(p):
        p.C *= 10.0
        p.D *= 10.0

    def rho_of_nacl(p):
        p.rho = 0.317

    def sigma_na(p):
        p.sigma[0] = 2 * 1.170
        p.sigma[1] = 1.170 + p.sigma[2] / 2

    def b_times_ten(p):
        p.b *= 10.0

    def cross_radius_not_a_sum(p):
        p.sigma[1] = p.sigma[0] / 2

    return [cd_unconverted, rho_of_nacl, sigma_na, b_times_ten, cross_radius_not_a_sum]


@pytest.mark.parametrize("mutate", _mutations(), ids=lambda f: f.__name__)
def test_each_plausible_mistake_fails_a_crystal_check(mutate):
    pp = Pot(mdgpu.potential("KCl"))
    mutate(pp)
    assert not crystal_ok("KCl", pp)


@pytest.mark.parametrize("salt,a", [("NaCl", 5.6178), ("KCl", 6.2254)])
@pytest.mark.parametrize("n", [3, 4])
def test_oracle_crystal_energy_equals_the_lattice_sum(salt, a, n):
    """The oracle's energy of the perfect crystal, E_real + E_recip/4 +
    E_self (analytic pairs, Ewald tuned to 1e-10), equals the lattice sum
    truncated at the same cutoff.  The cutoff sits 0.05 A inside L/2, away
    from every shell."""
    L = n * a
    rc = L / 2 - 0.05
    alpha, hm = mdgpu.ewald_params(rc, L, 1e-10)
    o = mdref.Oracle(n, n, n, a, alpha, hm, rc, 2.0, salt=salt, real_mode=1)
    o.set_state(f=np.zeros_like(o.r))
    o.calc_foc()
    b = o.b
    e_orc = (b.pot_real + b.pot_recip / 4 + b.pot_self) / (o.N / 2)
    e_sum = lattice_energy(Pot(mdgpu.potential(salt)), a, rmax=rc)
    assert abs(e_orc - e_sum) <= 1e-9 * abs(e_sum)
