"""Host-side mirror of libmdgpu_pr.so (include/prmdc_abi.h): pr-mdc's global
objects as ctypes structs, the library loader, and PrSystem, which owns one
pr-mdc system's arrays and binds them into a library's globals.

pr-mdc is the reference's constant-pressure program (pr-mdc/program): a
triclinic MD cell under the Parrinello-Rahman method.  libmdgpu_pr replaces
its real.c (real_space) and ewald.c (ewald1, ewald2, ewald3, calc_alpha).
The library fails loudly when it is missing; there is no CPU path here.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libmdgpu_pr.so")

PD = C.POINTER(C.c_double)
PPD = C.POINTER(PD)
PI = C.POINTER(C.c_int)

E_NODEVICE, E_HIP, E_ARG, E_NOMEM, E_STATE, E_RANGE = 1, 2, 3, 4, 6, 7

# pr-mdc/program/md.h:7-15 and init.c:67-123 (unit conversions)
SECD_2_FS = 4.074972637944947


class CalcControl(C.Structure):  # md.h:25-43
    _fields_ = [("skip_ewald", C.c_int), ("calc_max", C.c_int), ("delta_time_fs", C.c_double),
                ("kinds_of_ions", C.c_int), ("natoms_in_mol_unit", C.c_int),
                ("natoms_in_unit_cell", C.c_int), ("kinds_of_interaction", C.c_int)]


class CellParameters(C.Structure):  # md.h:45-85
    _fields_ = ([("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("vol", C.c_double)] +
                [(n, PD) for n in ("rx", "ry", "rz", "sx", "sy", "sz", "vx", "vy", "vz",
                                   "fx0", "fy0", "fz0", "fx", "fy", "fz")] +
                [(n, PPD) for n in ("hu", "hur", "hm", "hmr", "hmt", "hmtr", "hmv", "hmvt",
                                    "hmf", "g", "sigma")] +
                [(n, C.c_double) for n in ("Lx", "Ly", "Lz", "angle_ab", "angle_bc",
                                           "angle_ca")])


class SystemProperty(C.Structure):  # md.h:87-140
    _fields_ = [("ion", PI), ("N", C.c_int), ("step", C.c_int), ("dt", C.c_double),
                ("dt2", C.c_double), ("kB", C.c_double), ("kk", C.c_double), ("pot", C.c_double),
                ("kin", C.c_double), ("perMol", C.c_double), ("e2t", C.c_double),
                ("pp2gpa", C.c_double), ("a1", C.c_double), ("hm", C.c_int), ("hm_sqrt", C.c_int),
                ("radius", C.c_double), ("radius2", C.c_double), ("cutoff", C.c_double),
                ("cutoff2", C.c_double), ("rcsi", PD), ("rsni", PD), ("table_division", C.c_int),
                ("pe1r1", PD), ("fe1r1", PD), ("lookup", PI)]


class PressureTemperature(C.Structure):  # md.h:142-158
    _fields_ = ([("t_control_step", C.c_int), ("p_control_step", C.c_int)] +
                [(n, C.c_double) for n in ("pres_scalling_factor", "temp_set", "temp_set_initial",
                                           "pres_trace")] +
                [(n, PPD) for n in ("pres_set", "pres_tensor", "virial", "kin_tensor")])


class IonProperty(C.Structure):  # md.h:160-163
    _fields_ = [("m", PD), ("Z", PD)]


class PrSet(C.Structure):  # md.h:192-198 (Parrinello-Rahman cell mass, dt, velocity, forces)
    _fields_ = [("mass", C.c_double), ("dt", C.c_double), ("v", PPD), ("f", PPD), ("f0", PPD)]


GLOBALS = (("ctl", CalcControl), ("cell", CellParameters), ("sys", SystemProperty),
           ("pt", PressureTemperature), ("ion", IonProperty), ("pr", PrSet))


def copy_globals(src, dst):
    """Points library dst's pr-mdc globals at what library src's hold (the
    same arrays, not copies)."""
    for name, cls in GLOBALS:
        C.memmove(C.addressof(cls.in_dll(dst, name)), C.addressof(cls.in_dll(src, name)),
                  C.sizeof(cls))


class MdgPrError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what}: libmdgpu_pr error {code}")
        self.code = code


_lib = None


def lib():
    """libmdgpu_pr.so, loaded once; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MdgPrError(E_NODEVICE, f"{LIB_PATH} is not built (make -C {PKG})")
        L = C.CDLL(LIB_PATH)
        for n in ("real_space", "ewald1", "ewald2", "ewald3", "calc_alpha"):
            getattr(L, n).restype = None
            getattr(L, n).argtypes = []
        L.mdg_pr_set_error_mode.argtypes = [C.c_int]
        L.mdg_pr_set_error_mode.restype = C.c_int
        # errors surface as MdgPrError from the status, not as an exit
        L.mdg_pr_set_error_mode(1)
        L.mdg_pr_status.restype = C.c_int
        L.mdg_pr_clear_status.restype = None
        L.mdg_pr_set_splits.argtypes = [C.c_int, C.c_int]
        L.mdg_pr_set_splits.restype = C.c_int
        L.mdg_pr_profile.argtypes = [C.c_int, PD, PD]
        L.mdg_pr_profile.restype = C.c_int
        L.mdg_pr_run.argtypes = [C.c_int, PD]
        L.mdg_pr_run.restype = C.c_int
        _lib = L
    return _lib


def globals_of(L):
    """(ctl, cell, sys, pt, ion) of a library that defines pr-mdc's globals."""
    return (CalcControl.in_dll(L, "ctl"), CellParameters.in_dll(L, "cell"),
            SystemProperty.in_dll(L, "sys"), PressureTemperature.in_dll(L, "pt"),
            IonProperty.in_dll(L, "ion"))


class Mat3:
    """A 3x3 double matrix laid out as pr-mdc's dmat2d (init.c:240-253)."""

    def __init__(self, a=None):
        self.a = np.zeros((3, 3)) if a is None else np.array(a, dtype=np.float64).copy()
        self.rows = (PD * 3)(*[self.a[i].ctypes.data_as(PD) for i in range(3)])

    @property
    def ptr(self):
        return C.cast(self.rows, PPD)


def _dp(a):
    return a.ctypes.data_as(PD)


class PrSystem:
    """One pr-mdc system's state in numpy arrays.

    r, f: (3, N) positions and forces; typ: species; Z: charges per
    species; hm: cell matrix (rows a, b, c); pe, fe, lookup, td: pr-mdc's
    table.  bind(L) points library L's globals at these arrays (and the
    scalars at these values); read(L) copies back pot and the virial."""

    def __init__(self, r, typ, Z, hm, pe, fe, lookup, td, a1, cutoff, hmax, m=None):
        self.N = int(r.shape[1])
        self.r = np.ascontiguousarray(r, dtype=np.float64).copy()
        self.v = np.zeros_like(self.r)
        self.f = np.zeros_like(self.r)
        self.typ = np.ascontiguousarray(typ, dtype=np.int32).copy()
        self.Z = np.array(Z, dtype=np.float64)
        self.m = np.array(m if m is not None else np.ones(len(Z)), dtype=np.float64)
        self.kinds = len(Z)
        self.pe, self.fe = np.array(pe, dtype=np.float64), np.array(fe, dtype=np.float64)
        self.lookup = np.array(lookup, dtype=np.int32)
        self.td = int(td)
        self.a1, self.cutoff, self.hmax = float(a1), float(cutoff), int(hmax)
        self.hm = Mat3(hm)
        self.hmr = Mat3(np.linalg.inv(np.asarray(hm)))
        self.g = Mat3()
        self.virial = Mat3()
        self.vol = abs(float(np.linalg.det(np.asarray(hm))))
        self.pot = 0.0

    def set_cell_from(self, hmr, g, vol):
        """Takes H^-1, g and the volume as pr-mdc computes them
        (cell.c:148-175, 214-252), so that both sides fold identically."""
        self.hmr.a[:] = hmr
        self.g.a[:] = g
        self.vol = float(vol)

    def bind(self, L):
        ctl, cell, sys_, pt, ion = globals_of(L)
        ctl.kinds_of_ions = self.kinds
        ctl.kinds_of_interaction = self.kinds * (self.kinds + 1) // 2
        ctl.skip_ewald = 0
        cell.rx, cell.ry, cell.rz = _dp(self.r[0]), _dp(self.r[1]), _dp(self.r[2])
        cell.fx, cell.fy, cell.fz = _dp(self.f[0]), _dp(self.f[1]), _dp(self.f[2])
        cell.vx, cell.vy, cell.vz = _dp(self.v[0]), _dp(self.v[1]), _dp(self.v[2])
        cell.hm, cell.hmr, cell.g = self.hm.ptr, self.hmr.ptr, self.g.ptr
        cell.vol = self.vol
        sys_.ion = self.typ.ctypes.data_as(PI)
        sys_.N = self.N
        sys_.a1 = self.a1
        sys_.cutoff = self.cutoff
        sys_.cutoff2 = self.cutoff * self.cutoff
        sys_.hm = self.hmax
        sys_.hm_sqrt = int(np.sqrt(np.float32(self.hmax)))
        sys_.table_division = self.td
        sys_.pe1r1, sys_.fe1r1 = _dp(self.pe), _dp(self.fe)
        sys_.lookup = self.lookup.ctypes.data_as(PI)
        sys_.pot = self.pot
        pt.virial = self.virial.ptr
        ion.Z, ion.m = _dp(self.Z), _dp(self.m)

    def read(self, L):
        self.pot = globals_of(L)[2].pot
        return self.pot


def call(name, what=None):
    """Calls one routine of libmdgpu_pr and raises on a library error."""
    L = lib()
    getattr(L, name)()
    st = L.mdg_pr_status()
    if st != 0:
        L.mdg_pr_clear_status()
        raise MdgPrError(st, what or name)


def set_splits(pair, kspace):
    rc = lib().mdg_pr_set_splits(pair, kspace)
    if rc != 0:
        raise MdgPrError(rc, "mdg_pr_set_splits")


def profile(reps):
    """(ms per force evaluation, [real_space, ewald1, ewald2, ewald3] ms),
    device work only, inputs resident."""
    ms = C.c_double()
    per = (C.c_double * 4)()
    rc = lib().mdg_pr_profile(reps, C.byref(ms), per)
    if rc != 0:
        raise MdgPrError(rc, "mdg_pr_profile")
    return ms.value, list(per)


def run(nsteps):
    """mdg_pr_run: nsteps of pr-mdc's step loop resident on the device, on
    the system in the library's globals; returns the per-step records
    (nsteps, 13): pot, kin, pres_trace, vol, hm[3][3]."""
    rec = np.zeros((max(nsteps, 1), 13))
    rc = lib().mdg_pr_run(nsteps, rec.ctypes.data_as(PD))
    if rc != 0:
        lib().mdg_pr_clear_status()
        raise MdgPrError(rc, "mdg_pr_run")
    return rec[:nsteps]
