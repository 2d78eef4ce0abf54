"""ctypes access to the parity checker (oracle/liboracle.so) and, where it
was built, the compiled mdc reference harness (oracle/_ref/libmdcref.so).

Test infrastructure: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py only.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libmdcref.so")

KB = 1.380658e-5          # Boltzmann constant [1e-18 J / K]
KK = 2.307079556          # e^2 / (4 pi eps0) [1e-18 J A]
SECD_2_FS = 4.074972637944947

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OrcPotential(C.Structure):
    _fields_ = [("Z", C.c_double * 2), ("m", C.c_double * 2), ("b", C.c_double),
                ("rho", C.c_double), ("c", C.c_double * 3), ("sigma", C.c_double * 3),
                ("C", C.c_double * 3), ("D", C.c_double * 3)]


class OrcBox(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("Ax", C.c_double), ("Ay", C.c_double), ("Az", C.c_double),
        ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
        ("a1", C.c_double), ("p2a2", C.c_double), ("a3", C.c_double),
        ("hm", C.c_int), ("hm_sqrt", C.c_int),
        ("radius", C.c_double), ("radius2", C.c_double),
        ("delta_time_fs", C.c_double), ("dt", C.c_double), ("dt2", C.c_double),
        ("table_division", C.c_int),
        ("kB", C.c_double), ("kk", C.c_double), ("e2t", C.c_double),
        ("real_mode", C.c_int),
        ("pp", OrcPotential),
        ("lookup", C.c_int * 4),
        ("pe1r1", _dp), ("fe1r1", _dp),
        ("rx", _dp), ("ry", _dp), ("rz", _dp), ("vx", _dp), ("vy", _dp), ("vz", _dp),
        ("fx", _dp), ("fy", _dp), ("fz", _dp), ("fx0", _dp), ("fy0", _dp), ("fz0", _dp),
        ("ion", _ip),
        ("step", C.c_int),
        ("pot", C.c_double), ("pot_real", C.c_double), ("pot_recip", C.c_double),
        ("pot_self", C.c_double),
        ("virX", C.c_double), ("virY", C.c_double), ("virZ", C.c_double),
        ("kin", C.c_double), ("kinX", C.c_double), ("kinY", C.c_double), ("kinZ", C.c_double),
        ("presX", C.c_double), ("presY", C.c_double), ("presZ", C.c_double), ("pres", C.c_double),
    ]


class OrcObserver(C.Structure):
    _fields_ = [("kinds", C.c_int), ("msd_on", C.c_int)] + \
        [(n, _dp) for n in ("rx_old", "ry_old", "rz_old", "dx", "dy", "dz")] + \
        [("value", C.c_double * 4), ("number_of_ion", C.c_int * 4),
         ("Lx2", C.c_double), ("Ly2", C.c_double), ("Lz2", C.c_double),
         ("oldk", C.c_double), ("count", C.c_int), ("scalings", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(ORACLE_SO)
        bp = C.POINTER(OrcBox)
        for name in ("orc_unit_converter", "orc_set_loc", "orc_mk_table", "orc_clear_foc",
                     "orc_real_space", "orc_real_space_analytic", "orc_reciprocal_space",
                     "orc_reciprocal_space3", "orc_calc_foc", "orc_next_v", "orc_next_r",
                     "orc_next_rv_verlet", "orc_calc_kin", "orc_moment_correction",
                     "orc_calc_press"):
            getattr(L, name).argtypes = [bp]
            getattr(L, name).restype = None
        L.orc_nve.argtypes = [bp, C.c_int, _dp]
        L.orc_nve.restype = None
        L.orc_erfcc.argtypes = [C.c_double]
        L.orc_erfcc.restype = C.c_double
        L.orc_potential_nacl.argtypes = [C.c_double, C.POINTER(OrcPotential)]
        L.orc_potential_kcl.argtypes = [C.c_double, C.POINTER(OrcPotential)]
        L.orc_count_kvectors.argtypes = [C.c_int]
        L.orc_count_kvectors.restype = C.c_int
        op = C.POINTER(OrcObserver)
        L.orc_calc_msd.argtypes = [bp, op]
        L.orc_calc_msd.restype = None
        L.orc_control_temp.argtypes = [bp, op, C.c_int, C.c_double]
        L.orc_control_temp.restype = C.c_int
        L.orc_newton.argtypes = [bp, op, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                 C.c_int]
        L.orc_newton.restype = None
        _lib = L
    return _lib


def _ptr(a, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct))


class Oracle:
    """One box of the CPU restatement.  Arrays live in numpy and are
    passed by pointer, so self.r / self.v / self.f are live views."""

    def __init__(self, nx, ny, nz, a, alpha, hm, radius, dt_fs=2.0, salt="NaCl",
                 real_mode=0):
        L = lib()
        self.b = OrcBox()
        b = self.b
        b.nx, b.ny, b.nz = nx, ny, nz
        b.Ax = b.Ay = b.Az = a
        b.a1, b.hm, b.radius, b.delta_time_fs = alpha, hm, radius, dt_fs
        b.real_mode = real_mode
        L.orc_unit_converter(C.byref(b))
        if salt == "NaCl":
            L.orc_potential_nacl(b.kk, C.byref(b.pp))
        elif salt == "KCl":
            L.orc_potential_kcl(b.kk, C.byref(b.pp))
        else:
            raise ValueError(salt)
        n = b.N
        nb = 3 * (b.table_division + 1)
        self.pe = np.zeros(nb)
        self.fe = np.zeros(nb)
        self.r = np.zeros((3, n))
        self.v = np.zeros((3, n))
        self.f = np.zeros((3, n))
        self.f0 = np.zeros((3, n))
        self.ion = np.zeros(n, dtype=np.int32)
        b.pe1r1, b.fe1r1 = _ptr(self.pe), _ptr(self.fe)
        b.rx, b.ry, b.rz = (_ptr(self.r[k]) for k in range(3))
        b.vx, b.vy, b.vz = (_ptr(self.v[k]) for k in range(3))
        b.fx, b.fy, b.fz = (_ptr(self.f[k]) for k in range(3))
        b.fx0, b.fy0, b.fz0 = (_ptr(self.f0[k]) for k in range(3))
        b.ion = _ptr(self.ion, C.c_int)
        L.orc_set_loc(C.byref(b))
        L.orc_mk_table(C.byref(b))
        b.step = 1
        # calc_msd's arrays and control_temp's statics (zero, as mdc's calloc)
        self.msd_old = np.zeros((3, n))
        self.msd_d = np.zeros((3, n))
        self.obs = OrcObserver()
        o = self.obs
        o.kinds = 2
        o.rx_old, o.ry_old, o.rz_old = (_ptr(self.msd_old[k]) for k in range(3))
        o.dx, o.dy, o.dz = (_ptr(self.msd_d[k]) for k in range(3))

    @property
    def N(self):
        return self.b.N

    @property
    def L(self):
        return np.array([self.b.Lx, self.b.Ly, self.b.Lz])

    def set_state(self, r=None, v=None, f=None):
        if r is not None:
            self.r[...] = r
        if v is not None:
            self.v[...] = v
        if f is not None:
            self.f[...] = f

    def call(self, name):
        getattr(lib(), "orc_" + name)(C.byref(self.b))

    def calc_foc(self):
        self.call("calc_foc")

    def nve(self, steps, energies=False):
        e = np.zeros(steps) if energies else None
        lib().orc_nve(C.byref(self.b), steps, _ptr(e) if energies else None)
        return e

    def calc_msd(self):
        lib().orc_calc_msd(C.byref(self.b), C.byref(self.obs))

    def control_temp(self, d_step, temp):
        return lib().orc_control_temp(C.byref(self.b), C.byref(self.obs), d_step, temp)

    def newton(self, first_step, steps, t_control_step=0, temp=0.0, t_until=10000,
               moment_every=0, msd=False):
        """main.c:26-63 without I/O, steps first_step .. first_step+steps-1."""
        self.obs.msd_on = 1 if msd else 0
        lib().orc_newton(C.byref(self.b), C.byref(self.obs), first_step, steps,
                         t_control_step, temp, t_until, moment_every)

    def msd_values(self):
        return np.array(self.obs.value[:2])

    def scalars(self):
        b = self.b
        return dict(pot=b.pot, pot_real=b.pot_real, pot_recip=b.pot_recip, pot_self=b.pot_self,
                    virX=b.virX, virY=b.virY, virZ=b.virZ, kin=b.kin, kinX=b.kinX,
                    kinY=b.kinY, kinZ=b.kinZ, presX=b.presX, presY=b.presY,
                    presZ=b.presZ, pres=b.pres)

    def masses(self):
        return np.array([self.b.pp.m[0], self.b.pp.m[1]])[self.ion]

    def charges(self):
        return np.array([self.b.pp.Z[0], self.b.pp.Z[1]])[self.ion]


def maxwell_velocities(masses, temp_k, seed):
    """Seeded Maxwell-Boltzmann velocities [A / time unit] with the centre
    of mass momentum removed the way mdc's moment_correction does."""
    rng = np.random.default_rng(seed)
    m = np.asarray(masses, dtype=np.float64)
    sig = np.sqrt(KB * temp_k / m)
    v = rng.standard_normal((3, m.size)) * sig
    p = (m * v).sum(axis=1) / m.size
    v -= p[:, None] / m
    return v


def jitter_positions(r, L, amp, seed):
    rng = np.random.default_rng(seed)
    out = r + rng.uniform(-amp, amp, size=r.shape)
    L = np.asarray(L)[:, None]
    return np.mod(out, L)


# ---------------------------------------------------------------- reference
_refs = {}


def ref_available(material=None):
    return os.path.exists(ref_path(material))


def ref_path(material=None):
    """oracle/_ref/libmdcref.so (mdc/program), or the harness built with the
    material plugin of mdc/examples/<material> ('.' in the name as '_')."""
    if material is None:
        return REF_SO
    return os.path.join(ROOT, "oracle", "_ref", f"libmdcref_{material}.so")


def ref(material=None):
    path = ref_path(material)
    if path not in _refs:
        R = C.CDLL(path)
        R.ref_setup.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                C.c_double, C.c_double]
        R.ref_setup.restype = C.c_int
        for n in ("ref_set_positions", "ref_set_velocities", "ref_set_forces"):
            getattr(R, n).argtypes = [_dp, _dp, _dp]
        R.ref_get_state.argtypes = [_dp]
        R.ref_get_species.argtypes = [_ip]
        R.ref_get_scalars.argtypes = [_dp]
        R.ref_get_table.argtypes = [_dp, _dp]
        R.ref_erfcc.argtypes = [C.c_double]
        R.ref_erfcc.restype = C.c_double
        R.ref_run.argtypes = [C.c_int, C.c_int]
        R.ref_control_temp.argtypes = [C.c_int, C.c_double]
        R.ref_get_msd.argtypes = [_dp, _dp, _ip]
        R.ref_newton.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
                                 C.c_int]
        for n in ("ref_print_to_file", "ref_r_position", "ref_r_velocities"):
            getattr(R, n).argtypes = [C.c_char_p]
            getattr(R, n).restype = C.c_int
        R.ref_get_ions.argtypes = [_dp, _dp, _ip, _dp]
        _refs[path] = R
    return _refs[path]


class Reference:
    """The compiled mdc sources (one global system per process)."""

    def __init__(self, nx, ny, nz, a, alpha, hm, radius, dt_fs=2.0, material=None):
        self.R = ref(material)
        self.N = self.R.ref_setup(nx, ny, nz, a, alpha, hm, radius, dt_fs)

    def set_state(self, r=None, v=None, f=None):
        R = self.R
        if r is not None:
            r = np.ascontiguousarray(r, dtype=np.float64)
            R.ref_set_positions(_ptr(r[0]), _ptr(r[1]), _ptr(r[2]))
        if v is not None:
            v = np.ascontiguousarray(v, dtype=np.float64)
            R.ref_set_velocities(_ptr(v[0]), _ptr(v[1]), _ptr(v[2]))
        if f is not None:
            f = np.ascontiguousarray(f, dtype=np.float64)
            R.ref_set_forces(_ptr(f[0]), _ptr(f[1]), _ptr(f[2]))

    def state(self):
        out = np.zeros(12 * self.N)
        self.R.ref_get_state(_ptr(out))
        out = out.reshape(4, 3, self.N)
        return dict(r=out[0], v=out[1], f=out[2], f0=out[3])

    def species(self):
        out = np.zeros(self.N, dtype=np.int32)
        self.R.ref_get_species(_ptr(out, C.c_int))
        return out

    def scalars(self):
        out = np.zeros(12)
        self.R.ref_get_scalars(_ptr(out))
        keys = ("pot", "virX", "virY", "virZ", "kin", "kinX", "kinY", "kinZ",
                "presX", "presY", "presZ", "pres")
        return dict(zip(keys, out))

    def table(self):
        n = 3 * (self.R.ref_table_division() + 1)
        pe, fe = np.zeros(n), np.zeros(n)
        self.R.ref_get_table(_ptr(pe), _ptr(fe))
        return pe, fe

    def call(self, name):
        getattr(self.R, "ref_" + name)()

    def run(self, first_step, nsteps):
        self.R.ref_run(first_step, nsteps)

    def newton(self, first_step, nsteps, t_control_step=0, temp=0.0, t_until=10000,
               moment_every=0, msd=False):
        self.R.ref_newton(first_step, nsteps, t_control_step, temp, t_until, moment_every,
                         1 if msd else 0)

    def control_temp(self, d_step, temp):
        self.R.ref_control_temp(d_step, temp)

    def reset_control_temp(self):
        self.R.ref_reset_control_temp()

    def write(self, which, path):
        """files.c's print_to_file / r_position / r_velocities into path."""
        assert getattr(self.R, "ref_" + which)(os.fsencode(str(path))) == 0
        return open(path).read()

    def msd(self):
        """(old positions 3xN, summed displacements 3xN, value[2], number[2])"""
        arr = np.zeros(6 * self.N)
        val = np.zeros(4)
        num = np.zeros(4, dtype=np.int32)
        self.R.ref_get_msd(_ptr(arr), _ptr(val), _ptr(num, C.c_int))
        arr = arr.reshape(6, self.N)
        return arr[:3], arr[3:], val[:2], num[:2]

    def ions(self):
        """The material's species data: kinds, Z, m, lookup, L, ions per mol unit."""
        k = self.R.ref_kinds()
        Z, m, L = np.zeros(4), np.zeros(4), np.zeros(3)
        lk = np.zeros(16, dtype=np.int32)
        self.R.ref_get_ions(_ptr(Z), _ptr(m), _ptr(lk, C.c_int), _ptr(L))
        return dict(kinds=k, Z=Z[:k], m=m[:k], lookup=lk[:k * k], L=L,
                    mol_unit=self.R.ref_natoms_in_mol_unit())
