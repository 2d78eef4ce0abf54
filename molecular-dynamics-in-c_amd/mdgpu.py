"""ctypes binding of libmdgpu's C ABI (include/mdgpu.h, include/mdc_abi.h).

This is the Python side a test, benchmark or notebook uses; the product is
the C library.  Nothing here computes: every call goes to lib/libmdgpu.so,
and a missing library raises instead of falling back to anything else.
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libmdgpu.so")

OK, E_NODEVICE, E_HIP, E_ARG, E_NOMEM, E_BUDGET, E_STATE, E_RANGE, E_TIMEOUT = range(9)
FP64, MIXED = 0, 1
TABLE, ANALYTIC = 0, 1
DROPIN_EXIT, DROPIN_STATUS = 0, 1   # mdg_dropin_set_error_mode (mdc_abi.h)
CELLS_AUTO, ALL_PAIRS, CELL_LIST = -1, 0, 1
TERM_REAL, TERM_RECIP = 1, 2

SECD_2_FS = 4.074972637944947
KB = 1.380658e-5


class MdgError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what}: {errstr(code)} (code {code})")
        self.code = code


class Potential(C.Structure):
    _fields_ = [("Z", C.c_double * 2), ("m", C.c_double * 2), ("b", C.c_double),
                ("rho", C.c_double), ("c", C.c_double * 3), ("sigma", C.c_double * 3),
                ("C", C.c_double * 3), ("D", C.c_double * 3)]


class BoxParams(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("ax", C.c_double), ("ay", C.c_double), ("az", C.c_double),
                ("alpha", C.c_double), ("hm", C.c_int), ("rc", C.c_double),
                ("dt_fs", C.c_double), ("real_mode", C.c_int), ("cells", C.c_int),
                ("pot", Potential)]


class Scalars(C.Structure):
    _fields_ = [("pot", C.c_double), ("pot_real", C.c_double), ("pot_recip", C.c_double),
                ("pot_self", C.c_double), ("etot", C.c_double), ("vir", C.c_double * 3),
                ("kin", C.c_double), ("kin_xyz", C.c_double * 3), ("pres", C.c_double),
                ("pres_xyz", C.c_double * 3), ("step", C.c_longlong)]

    def as_dict(self):
        return dict(pot=self.pot, pot_real=self.pot_real, pot_recip=self.pot_recip,
                    pot_self=self.pot_self, etot=self.etot, vir=list(self.vir), kin=self.kin,
                    kin_xyz=list(self.kin_xyz), pres=self.pres, pres_xyz=list(self.pres_xyz),
                    step=self.step)


class Job(C.Structure):
    _fields_ = [("arrival_s", C.c_double), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("temp_k", C.c_double), ("steps", C.c_int), ("seed", C.c_ulonglong)]


class ServeConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_slots", C.c_int), ("precision", C.c_int),
                ("real_mode", C.c_int), ("eps", C.c_double), ("rc_max", C.c_double),
                ("dt_fs", C.c_double), ("one_at_a_time", C.c_int), ("realtime", C.c_int),
                ("steps_per_tick", C.c_int), ("auto_split", C.c_int)]


class JobResult(C.Structure):
    _fields_ = [("admit_s", C.c_double), ("finish_s", C.c_double), ("latency_s", C.c_double),
                ("etot_first", C.c_double), ("etot_last", C.c_double),
                ("final_checksum", C.c_double), ("status", C.c_int)]


class Tiling(C.Structure):
    _fields_ = [("cells", C.c_int), ("cell_w", C.c_double), ("rsplit", C.c_int),
                ("nchunk", C.c_int), ("ngroup", C.c_int), ("tuned", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class TuneReport(C.Structure):
    _fields_ = [("classes", C.c_int), ("runs", C.c_int), ("seconds", C.c_double)]


class ServeStats(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("ion_steps", C.c_double),
                ("ion_steps_per_s", C.c_double), ("p50_latency_s", C.c_double),
                ("p99_latency_s", C.c_double), ("mean_latency_s", C.c_double),
                ("max_resident", C.c_int), ("ticks", C.c_int), ("rejected", C.c_int),
                ("peak_bytes", C.c_size_t)]


# ---- mdc's globals (mdc/program/md.h:23-144, potential.h:2-28) ----------
_dp = C.POINTER(C.c_double)


class CalcControl(C.Structure):
    _fields_ = [("delta_time_fs", C.c_double), ("calc_max", C.c_int), ("temp", C.c_double),
                ("t_control_step", C.c_int), ("p_control_step", C.c_int),
                ("set_press_GPa_X", C.c_double), ("set_press_GPa_Y", C.c_double),
                ("set_press_GPa_Z", C.c_double), ("press_X", C.c_double),
                ("press_Y", C.c_double), ("press_Z", C.c_double), ("kinds_of_ions", C.c_int),
                ("natoms_in_mol_unit", C.c_int), ("natoms_in_unit_cell", C.c_int),
                ("kinds_of_interaction", C.c_int)]


class SystemProperty(C.Structure):
    _fields_ = [("rx", _dp), ("ry", _dp), ("rz", _dp), ("vx", _dp), ("vy", _dp), ("vz", _dp),
                ("fx0", _dp), ("fy0", _dp), ("fz0", _dp), ("fx", _dp), ("fy", _dp), ("fz", _dp),
                ("ion", C.POINTER(C.c_short)), ("N", C.c_short), ("nx", C.c_short),
                ("ny", C.c_short), ("nz", C.c_short),
                ("Ax", C.c_double), ("Ay", C.c_double), ("Az", C.c_double),
                ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
                ("kB", C.c_double), ("kk", C.c_double), ("perMol", C.c_double),
                ("e2t", C.c_double), ("e2tX", C.c_double), ("e2tY", C.c_double),
                ("e2tZ", C.c_double), ("pp2gpa", C.c_double), ("step", C.c_int),
                ("dt", C.c_double), ("dt2", C.c_double), ("a1", C.c_double),
                ("p2a2", C.c_double), ("a3", C.c_double), ("hm", C.c_short),
                ("hm_sqrt", C.c_short), ("radius", C.c_double), ("radius2", C.c_double),
                ("rcsi", _dp), ("rsni", _dp), ("table_division", C.c_int),
                ("pe1r1", _dp), ("fe1r1", _dp), ("lookup", C.POINTER(C.c_int)),
                ("virX", C.c_double), ("virY", C.c_double), ("virZ", C.c_double),
                ("presX", C.c_double), ("presY", C.c_double), ("presZ", C.c_double),
                ("pres", C.c_double), ("pres_scalling_factor", C.c_double),
                ("pot", C.c_double), ("kin", C.c_double), ("kinX", C.c_double),
                ("kinY", C.c_double), ("kinZ", C.c_double)]


class IonProperty(C.Structure):
    _fields_ = [("m", _dp), ("Z", _dp)]


class HugginsMayer(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "b", "c_na_na", "c_na_cl", "c_cl_cl", "sigma_na", "sigma_cl", "rho", "C_na_na",
        "C_na_cl", "C_cl_cl", "D_na_na", "D_na_cl", "D_cl_cl")]


class MsdVariables(C.Structure):
    """mdc/program/md.h:152-172"""
    _fields_ = [("number_of_ion", C.POINTER(C.c_int))] + \
        [(n, _dp) for n in ("rx_old", "ry_old", "rz_old", "dx", "dy", "dz", "value")] + \
        [(n, C.c_double) for n in ("Lx2", "Ly2", "Lz2")]


class Observe(C.Structure):
    """mdg_observe: the rest of mdc's step loop (main.c:44-60) per box."""
    _fields_ = [("msd", C.c_int), ("t_control_step", C.c_int), ("temp_K", C.c_double),
                ("t_control_until", C.c_longlong), ("moment_every", C.c_int)]


class OutFields(C.Structure):
    """mdg_out_fields: the values of one line of mdc's files/out."""
    _fields_ = [("step", C.c_int), ("kin", C.c_double), ("pot", C.c_double),
                ("pres", C.c_double), ("n_ions", C.c_int), ("natoms_in_mol_unit", C.c_int),
                ("L", C.c_double * 3), ("cells", C.c_int * 3), ("kinds", C.c_int),
                ("msd", C.c_double * 4)]


class Frame(C.Structure):
    _fields_ = [("box", C.c_int), ("n_ions", C.c_int), ("step", C.c_longlong),
                ("r", _dp), ("v", _dp), ("species", C.POINTER(C.c_int)),
                ("scalars", Scalars), ("msd", C.c_double * 4)]


FRAME_FN = C.CFUNCTYPE(None, C.POINTER(Frame), C.c_void_p)


class TableBox(C.Structure):
    """mdg_table_box: a box of one of mdc's tabulated materials."""
    _fields_ = [("n_ions", C.c_int), ("L", C.c_double * 3), ("cells", C.c_int * 3),
                ("natoms_in_mol_unit", C.c_int), ("kinds", C.c_int),
                ("Z", C.c_double * 4), ("m", C.c_double * 4), ("lookup", C.c_int * 16),
                ("ntypes", C.c_int), ("table_division", C.c_int), ("pe", _dp), ("fe", _dp),
                ("alpha", C.c_double), ("hm", C.c_int), ("rc", C.c_double),
                ("dt_fs", C.c_double), ("cells_mode", C.c_int)]


def table_box(n_ions, L, cells, mol_unit, Z, m, lookup, pe, fe, alpha, hm, rc, dt_fs=2.0,
              cells_mode=CELLS_AUTO):
    """A TableBox from mdc's per-material data (the numpy tables are kept
    alive on the returned object until the batch has copied them)."""
    k = len(Z)
    tb = TableBox()
    tb.n_ions, tb.kinds, tb.ntypes = n_ions, k, k * (k + 1) // 2
    tb.L[:] = list(L)
    tb.cells[:] = list(cells)
    tb.natoms_in_mol_unit = mol_unit
    for q in range(k):
        tb.Z[q], tb.m[q] = Z[q], m[q]
    for q, t in enumerate(lookup):
        tb.lookup[q] = int(t)
    pe = np.ascontiguousarray(pe, dtype=np.float64)
    fe = np.ascontiguousarray(fe, dtype=np.float64)
    tb.table_division = pe.size // tb.ntypes - 1
    tb.pe, tb.fe = _p(pe), _p(fe)
    tb._keep = (pe, fe)
    tb.alpha, tb.hm, tb.rc, tb.dt_fs, tb.cells_mode = alpha, hm, rc, dt_fs, cells_mode
    return tb
WRITE_POSITIONS, WRITE_VELOCITIES = 1, 2


DROPIN_ROUTINES = ("calc_foc", "clear_foc", "real_space", "reciprocal_space",
                   "reciprocal_space3", "next_rv_verlet", "next_r", "next_v", "calc_kin",
                   "moment_correction", "calc_press", "mk_table")

_lib = None


def lib():
    """Load lib/libmdgpu.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: build it with __graft_entry__.build() "
                               "or `make -C molecular-dynamics-in-c_amd`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        sig = {
            "mdg_device_count": ([C.POINTER(C.c_int)], C.c_int),
            "mdg_potential_nacl": ([C.POINTER(Potential)], C.c_int),
            "mdg_potential_kcl": ([C.POINTER(Potential)], C.c_int),
            "mdg_rocksalt": ([C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                              _dp, C.POINTER(C.c_int)], C.c_int),
            "mdg_maxwell": ([C.c_ulonglong, C.c_double, C.c_int, C.POINTER(C.c_int),
                             C.POINTER(Potential), _dp], C.c_int),
            "mdg_ewald_params": ([C.c_double, C.c_double, C.c_double, _dp,
                                  C.POINTER(C.c_int)], C.c_int),
            "mdg_batch_create": ([C.c_int, C.POINTER(BoxParams), C.c_int, C.POINTER(vp)],
                                 C.c_int),
            "mdg_batch_destroy": ([vp], C.c_int),
            "mdg_batch_create_tabulated": ([C.c_int, C.POINTER(TableBox), C.c_int,
                                            C.POINTER(vp)], C.c_int),
            "mdg_batch_n_ions": ([vp, C.c_int], C.c_int),
            "mdg_batch_box_bytes": ([vp, C.c_int], C.c_size_t),
            "mdg_batch_set_state": ([vp, C.c_int, _dp, _dp, C.POINTER(C.c_int)], C.c_int),
            "mdg_batch_get_state": ([vp, C.c_int, _dp, _dp, _dp, _dp], C.c_int),
            "mdg_batch_forces": ([vp], C.c_int),
            "mdg_batch_force_terms": ([vp, C.c_int], C.c_int),
            "mdg_batch_step": ([vp, C.c_int], C.c_int),
            "mdg_batch_prepare": ([vp], C.c_int),
            "mdg_batch_tune_schedule": ([vp, C.c_int, _dp], C.c_int),
            "mdg_batch_scalars": ([vp, C.c_int, C.POINTER(Scalars)], C.c_int),
            "mdg_batch_sync": ([vp], C.c_int),
            "mdg_batch_profile": ([vp, C.c_int, C.c_int, C.c_void_p, _dp,
                                   C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
            "mdg_batch_work": ([vp, C.c_int, C.POINTER(C.c_longlong), C.POINTER(C.c_int),
                                C.POINTER(C.c_int)], C.c_int),
            "mdg_error_string": ([C.c_int], C.c_char_p),
            "mdg_serve": ([C.POINTER(Job), C.c_int, C.POINTER(ServeConfig),
                           C.POINTER(JobResult), C.POINTER(ServeStats)], C.c_int),
            "mdg_tune": ([C.c_int, C.POINTER(BoxParams), C.c_int, C.c_longlong,
                          C.POINTER(TuneReport)], C.c_int),
            "mdg_tune_clear": ([], C.c_int),
            "mdg_copy_probe": ([C.c_size_t, C.c_int, C.POINTER(C.c_double)], C.c_int),
            "mdg_ewald_rc_rule": ([C.c_int, C.c_double, C.c_double, C.POINTER(C.c_double)],
                                  C.c_int),
            "mdg_serve_job_params": ([C.POINTER(Job), C.POINTER(ServeConfig),
                                      C.POINTER(BoxParams)], C.c_int),
            "mdg_tune_save": ([C.c_char_p], C.c_int),
            "mdg_tune_load": ([C.c_char_p], C.c_int),
            "mdg_box_tiling": ([C.POINTER(BoxParams), C.c_int, C.POINTER(Tiling)], C.c_int),
            "mdg_forces_noalloc": ([C.POINTER(BoxParams), C.c_int, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
            "mdg_dropin_set_error_mode": ([C.c_int], C.c_int),
            "mdg_dropin_status": ([], C.c_int),
            "mdg_dropin_clear_status": ([], None),
            "mdg_dropin_release": ([], None),
            "erfcc": ([C.c_double], C.c_double),
            "mdg_batch_set_observe": ([vp, C.c_int, C.POINTER(Observe)], C.c_int),
            "mdg_batch_observed": ([vp, C.c_int, _dp, C.POINTER(C.c_longlong)], C.c_int),
            "mdg_batch_msd_state": ([vp, C.c_int, _dp, _dp], C.c_int),
            "calc_msd": ([], None),
            "mdg_batch_step_recorded": ([vp, C.c_int, C.c_int, FRAME_FN, C.c_void_p], C.c_int),
            "mdg_batch_write_mdc_files": ([vp, C.c_int, C.c_int, C.c_int, C.c_char_p], C.c_int),
            "mdg_format_out": ([C.c_char_p, C.c_size_t, C.POINTER(OutFields)], C.c_int),
            "mdg_format_positions": ([C.c_char_p, C.c_size_t, C.c_int, C.POINTER(C.c_int),
                                      _dp], C.c_int),
            "mdg_format_velocities": ([C.c_char_p, C.c_size_t, C.c_int, C.POINTER(C.c_int),
                                       _dp], C.c_int),
            "control_temp": ([C.c_int, C.c_double], None),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        for name in DROPIN_ROUTINES:
            f = getattr(L, name)
            f.argtypes = []
            f.restype = None
        # this mirror raises MdgError from the drop-in's status (dropin_call)
        # instead of letting a failing routine exit the interpreter
        L.mdg_dropin_set_error_mode(DROPIN_STATUS)
        _lib = L
    return _lib


def errstr(code):
    try:
        return lib().mdg_error_string(code).decode()
    except Exception:  # library missing
        return f"error {code}"


def check(code, what):
    if code != OK:
        raise MdgError(code, what)


def device_count():
    n = C.c_int(0)
    rc = lib().mdg_device_count(C.byref(n))
    return n.value if rc == OK else 0


def potential(salt="NaCl"):
    p = Potential()
    fn = {"NaCl": lib().mdg_potential_nacl, "KCl": lib().mdg_potential_kcl}[salt]
    check(fn(C.byref(p)), "potential")
    return p


def ewald_params(rc, l_max, eps):
    a, h = C.c_double(), C.c_int()
    check(lib().mdg_ewald_params(rc, l_max, eps, C.byref(a), C.byref(h)), "ewald_params")
    return a.value, h.value


def ewald_rc_rule(n_ions, l_min, rc_cap):
    """The library's real-space cutoff rule (mdg_ewald_rc_rule)."""
    rc = C.c_double()
    check(lib().mdg_ewald_rc_rule(n_ions, l_min, rc_cap, C.byref(rc)), "ewald_rc_rule")
    return rc.value


def rocksalt(nx, ny, nz, a):
    ax, ay, az = (a, a, a) if np.isscalar(a) else a
    n = 8 * nx * ny * nz
    r = np.zeros(3 * n)
    sp = np.zeros(n, dtype=np.int32)
    check(lib().mdg_rocksalt(nx, ny, nz, ax, ay, az, _p(r), _ip(sp)), "rocksalt")
    return r.reshape(3, n), sp


def maxwell(seed, temp_k, species, pot):
    species = np.ascontiguousarray(species, dtype=np.int32)
    v = np.zeros(3 * species.size)
    check(lib().mdg_maxwell(seed, temp_k, species.size, _ip(species), C.byref(pot), _p(v)),
          "maxwell")
    return v.reshape(3, species.size)


def box_params(nx, ny, nz, a, alpha, hm, rc, dt_fs=2.0, real_mode=TABLE, cells=CELLS_AUTO,
               salt="NaCl"):
    ax, ay, az = (a, a, a) if np.isscalar(a) else a
    p = BoxParams()
    p.nx, p.ny, p.nz = nx, ny, nz
    p.ax, p.ay, p.az = ax, ay, az
    p.alpha, p.hm, p.rc, p.dt_fs = alpha, hm, rc, dt_fs
    p.real_mode, p.cells = real_mode, cells
    p.pot = potential(salt) if isinstance(salt, str) else salt
    return p


def _p(a):
    return a.ctypes.data_as(_dp)


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class Batch:
    """A set of device-resident boxes stepped together."""

    def __init__(self, params, precision=FP64):
        arr = (BoxParams * len(params))(*params)
        h = C.c_void_p()
        check(lib().mdg_batch_create(len(params), arr, precision, C.byref(h)), "batch_create")
        self.h = h
        self.nbox = len(params)
        self.precision = precision

    @classmethod
    def tabulated(cls, boxes, precision=FP64):
        """Boxes of mdc's tabulated materials (TableBox, see table_box); their
        positions and species come from set_state."""
        arr = (TableBox * len(boxes))(*boxes)
        h = C.c_void_p()
        check(lib().mdg_batch_create_tabulated(len(boxes), arr, precision, C.byref(h)),
              "batch_create_tabulated")
        self = cls.__new__(cls)
        self.h, self.nbox, self.precision = h, len(boxes), precision
        return self

    def close(self):
        if self.h:
            lib().mdg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def n_ions(self, b=0):
        return lib().mdg_batch_n_ions(self.h, b)

    def box_bytes(self, b=0):
        return lib().mdg_batch_box_bytes(self.h, b)

    def set_state(self, b=0, r=None, v=None, species=None):
        n = self.n_ions(b)
        ra = None if r is None else np.ascontiguousarray(np.asarray(r, dtype=np.float64).reshape(3 * n))
        va = None if v is None else np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(3 * n))
        sa = None if species is None else np.ascontiguousarray(species, dtype=np.int32)
        check(lib().mdg_batch_set_state(self.h, b, None if ra is None else _p(ra),
                                        None if va is None else _p(va),
                                        None if sa is None else _ip(sa)), "set_state")

    def state(self, b=0):
        n = self.n_ions(b)
        out = [np.zeros(3 * n) for _ in range(4)]
        check(lib().mdg_batch_get_state(self.h, b, *(_p(x) for x in out)), "get_state")
        r, v, f, f0 = (x.reshape(3, n) for x in out)
        return dict(r=r, v=v, f=f, f0=f0)

    def forces(self):
        check(lib().mdg_batch_forces(self.h), "forces")

    def force_terms(self, terms):
        """f0 <- f, f <- the selected terms (TERM_REAL | TERM_RECIP)."""
        check(lib().mdg_batch_force_terms(self.h, terms), "force_terms")

    def step(self, n):
        check(lib().mdg_batch_step(self.h, n), "step")

    def prepare(self):
        """Capture and upload the step graph without stepping."""
        check(lib().mdg_batch_prepare(self.h), "prepare")

    def tune_schedule(self, nsteps=32):
        """Pick the step schedule's group split by timing (state restored);
        returns the chosen share."""
        share = C.c_double()
        check(lib().mdg_batch_tune_schedule(self.h, nsteps, C.byref(share)), "tune_schedule")
        return share.value

    def sync(self):
        check(lib().mdg_batch_sync(self.h), "sync")

    def scalars(self, b=0):
        s = Scalars()
        check(lib().mdg_batch_scalars(self.h, b, C.byref(s)), "scalars")
        return s.as_dict()

    def set_observe(self, b=0, msd=False, t_control_step=0, temp_K=0.0,
                    t_control_until=10000, moment_every=0):
        """Arm mdc's calc_msd / control_temp / moment_correction cadence for
        box b inside every step (main.c:44-60); resets their state."""
        o = Observe(1 if msd else 0, t_control_step, temp_K, t_control_until, moment_every)
        check(lib().mdg_batch_set_observe(self.h, b, C.byref(o)), "set_observe")

    def observed(self, b=0):
        msd = np.zeros(4)
        n = C.c_longlong()
        check(lib().mdg_batch_observed(self.h, b, _p(msd), C.byref(n)), "observed")
        return dict(msd=msd, scalings=n.value)

    def msd_state(self, b=0):
        n = self.n_ions(b)
        old, d = np.zeros(3 * n), np.zeros(3 * n)
        check(lib().mdg_batch_msd_state(self.h, b, _p(old), _p(d)), "msd_state")
        return old.reshape(3, n), d.reshape(3, n)

    def step_recorded(self, nsteps, every, keep=True):
        """Step nsteps, reading every box back after each `every` steps while
        the device goes on stepping.  keep: return the frames as a list of
        dicts (copies); else the frames are copied to the host and dropped."""
        frames = []

        def cb(fp, _user):
            f = fp.contents
            n = f.n_ions
            frames.append(dict(box=f.box, step=f.step,
                               r=np.ctypeslib.as_array(f.r, (3 * n,)).reshape(3, n).copy(),
                               v=np.ctypeslib.as_array(f.v, (3 * n,)).reshape(3, n).copy(),
                               scalars=f.scalars.as_dict(), msd=np.array(f.msd[:])))

        fn = FRAME_FN(cb) if keep else C.cast(None, FRAME_FN)
        check(lib().mdg_batch_step_recorded(self.h, nsteps, every, fn, None), "step_recorded")
        return frames

    def write_mdc_files(self, nsteps, every, what, directory):
        """mdc's files/out (and positions / velocities by `what`) per box
        under directory/box<k>/, one record per `every` steps."""
        check(lib().mdg_batch_write_mdc_files(self.h, nsteps, every, what,
                                              os.fsencode(directory)), "write_mdc_files")

    def work(self, b=0):
        p, k, n = C.c_longlong(), C.c_int(), C.c_int()
        check(lib().mdg_batch_work(self.h, b, C.byref(p), C.byref(k), C.byref(n)), "work")
        return dict(pairs=p.value, nk=k.value, n_ions=n.value)

    def profile(self, nsteps, alone=False):
        """Per kernel name: (ms per step over all its launches, launches per
        step), timed in the step's own schedule, or with every kernel
        launched alone over all boxes (alone=True)."""
        if alone:
            nsteps = -nsteps
        cap = 16
        names = (C.c_char * 32 * cap)()
        ms = np.zeros(cap)
        cnt = np.zeros(cap, dtype=np.int32)
        n = C.c_int()
        check(lib().mdg_batch_profile(self.h, nsteps, cap, names, _p(ms), _ip(cnt), C.byref(n)),
              "profile")
        return {bytes(names[k]).split(b"\0")[0].decode(): (float(ms[k]), int(cnt[k]))
                for k in range(n.value)}


def tune(boxes, precision, context_ions=0):
    """Autotune the tilings of every box class among `boxes` (BoxParams);
    returns the report as a dict."""
    arr = (BoxParams * len(boxes))(*boxes)
    rep = TuneReport()
    check(lib().mdg_tune(len(boxes), arr, precision, context_ions, C.byref(rep)), "tune")
    return {"classes": rep.classes, "runs": rep.runs, "seconds": rep.seconds}


def copy_probe(nbytes, reps=10):
    """Average ms of a plain device copy of nbytes (read nbytes + write nbytes)."""
    ms = C.c_double()
    check(lib().mdg_copy_probe(nbytes, reps, C.byref(ms)), "copy_probe")
    return ms.value


def tune_clear():
    check(lib().mdg_tune_clear(), "tune_clear")


def tune_save(path):
    check(lib().mdg_tune_save(os.fsencode(path)), "tune_save")


def tune_load(path):
    check(lib().mdg_tune_load(os.fsencode(path)), "tune_load")


def box_tiling(p, precision):
    t = Tiling()
    check(lib().mdg_box_tiling(C.byref(p), precision, C.byref(t)), "box_tiling")
    return t.as_dict()


def serve_job_params(job, cfg):
    p = BoxParams()
    check(lib().mdg_serve_job_params(C.byref(job), C.byref(cfg), C.byref(p)), "serve_job_params")
    return p


def serve(jobs, cfg):
    arr = (Job * len(jobs))(*jobs)
    res = (JobResult * len(jobs))()
    st = ServeStats()
    check(lib().mdg_serve(arr, len(jobs), C.byref(cfg), res, C.byref(st)), "serve")
    return list(res), st


def _text(fn, *args):
    n = fn(None, 0, *args)
    if n < 0:
        raise MdgError(n, "format")
    buf = C.create_string_buffer(n + 1)
    fn(buf, n + 1, *args)
    return buf.raw[:n].decode()


def format_out(step, kin, pot, pres, n_ions, L, cells, msd=(0.0, 0.0), natoms_in_mol_unit=2):
    """One line of mdc's files/out (files.c:44-65)."""
    f = OutFields(step, kin, pot, pres, n_ions, natoms_in_mol_unit, (C.c_double * 3)(*L),
                  (C.c_int * 3)(*cells), len(msd), (C.c_double * 4)(*msd))
    return _text(lib().mdg_format_out, C.byref(f))


def format_positions(species, r):
    """One r_position record (files.c:142-158); r: 3 x N."""
    sp = np.ascontiguousarray(species, dtype=np.int32)
    r = np.ascontiguousarray(r, dtype=np.float64)
    return _text(lib().mdg_format_positions, sp.size, _ip(sp), _p(r))


def format_velocities(species, v):
    """One r_velocities record (files.c:161-181); v: 3 x N [A / time unit]."""
    sp = np.ascontiguousarray(species, dtype=np.int32)
    v = np.ascontiguousarray(v, dtype=np.float64)
    return _text(lib().mdg_format_velocities, sp.size, _ip(sp), _p(v))


# ---- the drop-in surface -------------------------------------------------
def dropin_globals():
    L = lib()
    return (SystemProperty.in_dll(L, "sys"), IonProperty.in_dll(L, "ion"),
            CalcControl.in_dll(L, "ctl"), HugginsMayer.in_dll(L, "hm"))


def dropin_msd():
    return MsdVariables.in_dll(lib(), "msd")


def dropin_control_temp(d_step, temp):
    lib().control_temp(d_step, temp)
    st = lib().mdg_dropin_status()
    if st != OK:
        lib().mdg_dropin_clear_status()
        raise MdgError(st, "control_temp")


def dropin_call(name):
    getattr(lib(), name)()
    st = lib().mdg_dropin_status()
    if st != OK:
        lib().mdg_dropin_clear_status()
        raise MdgError(st, name)
