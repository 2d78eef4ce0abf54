"""Shared helpers of the GPU tests: wiring mdc's globals of the drop-in to
numpy arrays, and error measures."""
import ctypes as C

import numpy as np

import mdgpu
import mdref

A_NACL = 5.64
A_KCL = 6.29


def rel_err(a, b):
    """max |a - b| / max |b| over all components (norm-wise)."""
    d = np.abs(np.asarray(a) - np.asarray(b)).max()
    s = np.abs(np.asarray(b)).max()
    return d / s if s > 0 else d


def rms_rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum())


def cfg1_oracle(**kw):
    n = kw.pop("n", 3)
    a = kw.pop("a", A_NACL)
    L = n * a
    args = dict(alpha=5.6 / L, hm=27, radius=L / 2, dt_fs=2.0)
    args.update(kw)
    return mdref.Oracle(n, n, n, a, args["alpha"], args["hm"], args["radius"], args["dt_fs"],
                        salt=args.get("salt", "NaCl"), real_mode=args.get("real_mode", 0))


class DropIn:
    """mdc's globals inside libmdgpu, pointed at numpy arrays mirroring an
    oracle box (the caller's arrays stay authoritative, as in mdc)."""

    def __init__(self, o):
        self.o = o
        self.sys, self.ion, self.ctl, self.hm = mdgpu.dropin_globals()
        n = o.N
        self.r = np.zeros((3, n))
        self.v = np.zeros((3, n))
        self.f = np.zeros((3, n))
        self.f0 = np.zeros((3, n))
        self.species = o.ion.astype(np.int16)
        self.pe = o.pe.copy()
        self.fe = o.fe.copy()
        self.lookup = np.array([0, 1, 1, 2], dtype=np.int32)
        self.Z = np.array(list(o.b.pp.Z))
        self.m = np.array(list(o.b.pp.m))
        s, b = self.sys, o.b
        dp = C.POINTER(C.c_double)
        for name, arr in (("rx", self.r[0]), ("ry", self.r[1]), ("rz", self.r[2]),
                          ("vx", self.v[0]), ("vy", self.v[1]), ("vz", self.v[2]),
                          ("fx", self.f[0]), ("fy", self.f[1]), ("fz", self.f[2]),
                          ("fx0", self.f0[0]), ("fy0", self.f0[1]), ("fz0", self.f0[2]),
                          ("pe1r1", self.pe), ("fe1r1", self.fe)):
            setattr(s, name, arr.ctypes.data_as(dp))
        s.ion = self.species.ctypes.data_as(C.POINTER(C.c_short))
        s.lookup = self.lookup.ctypes.data_as(C.POINTER(C.c_int))
        s.N = n
        s.nx, s.ny, s.nz = b.nx, b.ny, b.nz
        s.Lx, s.Ly, s.Lz = b.Lx, b.Ly, b.Lz
        s.a1, s.a3, s.p2a2 = b.a1, b.a3, b.p2a2
        s.hm, s.hm_sqrt = b.hm, b.hm_sqrt
        s.radius, s.radius2 = b.radius, b.radius2
        s.table_division = b.table_division
        s.dt, s.dt2 = b.dt, b.dt2
        self.ion.Z = self.Z.ctypes.data_as(dp)
        self.ion.m = self.m.ctypes.data_as(dp)
        self.ctl.kinds_of_ions = 2
        self.ctl.kinds_of_interaction = 3

    def set_state(self, r=None, v=None, f=None, f0=None):
        for dst, src in ((self.r, r), (self.v, v), (self.f, f), (self.f0, f0)):
            if src is not None:
                dst[...] = src

    def call(self, name):
        mdgpu.dropin_call(name)

    def release(self):
        mdgpu.lib().mdg_dropin_release()
        s = self.sys
        for name in ("rx", "ry", "rz", "vx", "vy", "vz", "fx", "fy", "fz", "fx0", "fy0", "fz0",
                     "pe1r1", "fe1r1", "ion", "lookup"):
            setattr(s, name, None)
        self.ion.Z = self.ion.m = None
