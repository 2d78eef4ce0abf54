"""The C-ABI library: loads, exports what include/*.h declares, lays out
mdc's globals as mdc does, and its host-side setup matches the oracle.
No device computation here (these run without a GPU)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import mdgpu
import mdref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
MDC_PROGRAM = "/root/reference/mdc/program"
PRMDC_PROGRAM = "/root/reference/pr-mdc/program"


def declared_symbols():
    names = set()
    for h in ("mdgpu.h", "mdc_abi.h"):
        src = open(os.path.join(INC, h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?(?:int|void|double|size_t|char)\s*\*?\s*(\w+)\s*\(",
                             src, flags=re.M):
            names.add(m.group(1))
        for m in re.finditer(r"^extern\s+\w+\s+(\w+);", src, flags=re.M):
            names.add(m.group(1))
    return names


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", mdgpu.LIB_PATH], check=True,
                         capture_output=True, text=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if ln.strip()}


def test_library_exports_every_declared_symbol():
    decl = declared_symbols()
    assert {"calc_foc", "real_space", "reciprocal_space", "next_rv_verlet", "sys",
            "mdg_batch_step", "mdg_serve"} <= decl
    exp = exported_symbols()
    assert decl <= exp, sorted(decl - exp)
    # nothing but the C ABI leaks out of the library
    assert exp <= decl, sorted(exp - decl)
    mdgpu.lib()   # loads with all its dependencies


def _layout(tmp_path, mdc_headers):
    exe = tmp_path / ("probe_mdc" if mdc_headers else "probe_abi")
    cmd = ["gcc", "-std=c99", "-fcommon", "-o", str(exe), os.path.join(ROOT, "tests", "abi_layout.c")]
    cmd += ["-DUSE_MDC_HEADERS", "-I" + MDC_PROGRAM] if mdc_headers else ["-I" + INC]
    subprocess.run(cmd, check=True, capture_output=True)
    return subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout


def test_ctypes_mirror_matches_header_layout(tmp_path):
    lines = dict(ln.rsplit(" ", 1) for ln in _layout(tmp_path, False).splitlines())
    for cname, cls in (("system_property", mdgpu.SystemProperty),
                       ("calc_control", mdgpu.CalcControl),
                       ("ion_property", mdgpu.IonProperty),
                       ("huggins_mayer_potential_set", mdgpu.HugginsMayer),
                       ("msd_calculation_variables", mdgpu.MsdVariables)):
        assert int(lines[cname]) == C.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert int(lines[f"{cname}.{f}"]) == getattr(cls, f).offset, (cname, f)


@pytest.mark.skipif(not os.path.isdir(MDC_PROGRAM), reason="reference checkout absent")
def test_header_layout_is_mdc_layout(tmp_path):
    assert _layout(tmp_path, False) == _layout(tmp_path, True)


def test_host_setup_matches_oracle():
    L = mdgpu.lib()
    for x in (0.05, 0.3, 1.1, 2.7, 4.9):
        assert L.erfcc(x) == mdref.lib().orc_erfcc(x)
    o = mdref.Oracle(3, 2, 4, 5.64, 0.3, 27, 5.0)
    r, sp = mdgpu.rocksalt(3, 2, 4, 5.64)
    assert np.array_equal(r, o.r) and np.array_equal(sp, o.ion)
    p = mdgpu.potential("NaCl")
    for f in ("Z", "m", "c", "sigma", "C", "D"):
        assert list(getattr(p, f)) == list(getattr(o.b.pp, f)), f
    assert (p.b, p.rho) == (o.b.pp.b, o.b.pp.rho)
    k = mdgpu.potential("KCl")
    ok = mdref.Oracle(2, 2, 2, 6.29, 0.3, 27, 5.0, salt="KCl")
    for f in ("Z", "m", "c", "sigma", "C", "D"):
        assert list(getattr(k, f)) == list(getattr(ok.b.pp, f)), f


def test_dropin_mk_table_matches_oracle():
    """mk_table through mdc's globals (host routine of the drop-in)."""
    sysg, iong, ctlg, hmg = mdgpu.dropin_globals()
    o = mdref.Oracle(3, 3, 3, 5.64, 5.6 / 16.92, 27, 8.46)
    nb = o.b.table_division + 1
    pe, fe = np.zeros(3 * nb), np.zeros(3 * nb)
    lookup = np.zeros(4, dtype=np.int32)
    Z = np.array(list(o.b.pp.Z))
    m = np.array(list(o.b.pp.m))
    sysg.pe1r1 = pe.ctypes.data_as(C.POINTER(C.c_double))
    sysg.fe1r1 = fe.ctypes.data_as(C.POINTER(C.c_double))
    sysg.lookup = lookup.ctypes.data_as(C.POINTER(C.c_int))
    sysg.table_division = o.b.table_division
    sysg.a1, sysg.a3 = o.b.a1, o.b.a3
    iong.Z = Z.ctypes.data_as(C.POINTER(C.c_double))
    iong.m = m.ctypes.data_as(C.POINTER(C.c_double))
    ctlg.kinds_of_ions = 2
    hmg.b, hmg.rho = 3.38 / 100.0, 0.317
    hmg.c_na_na, hmg.c_na_cl, hmg.c_cl_cl = 1.25, 1.0, 0.75
    hmg.sigma_na, hmg.sigma_cl = 1.170, 1.585
    hmg.C_na_na, hmg.C_na_cl, hmg.C_cl_cl = 1.68 / 10.0, 11.2 / 10.0, 116.0 / 10.0
    hmg.D_na_na, hmg.D_na_cl, hmg.D_cl_cl = 0.8 / 10.0, 13.9 / 10.0, 233.0 / 10.0
    mdgpu.dropin_call("mk_table")
    assert np.array_equal(pe, o.pe) and np.array_equal(fe, o.fe)
    assert list(lookup) == [0, 1, 1, 2]
    sysg.pe1r1 = sysg.fe1r1 = None
    sysg.lookup = None
    iong.Z = iong.m = None


def test_ewald_params_rule():
    a, hm = mdgpu.ewald_params(10.0, 45.12, 1e-6)
    s = np.sqrt(np.log(1e6))
    assert a == pytest.approx(s / 10.0)
    assert hm == int(np.floor((s * a * 45.12 / np.pi) ** 2))
    assert 380 < hm < 400   # cfg2: alpha ~ 0.372, hm ~ 393


def test_ewald_rc_rule():
    """mdg_ewald_rc_rule: 16.5 A (N/32768)^(1/6), at least 10 A, at most
    rc_cap and L_min/2; argument errors for empty boxes."""
    assert mdgpu.ewald_rc_rule(32768, 100.64, 30.0) == pytest.approx(16.5)
    assert mdgpu.ewald_rc_rule(4096, 45.12, 30.0) == pytest.approx(16.5 / 2 ** 0.5)
    assert mdgpu.ewald_rc_rule(8 * 32768, 500.0, 30.0) == pytest.approx(16.5 * 2 ** 0.5)
    assert mdgpu.ewald_rc_rule(512, 22.56, 30.0) == pytest.approx(10.0)      # floor
    assert mdgpu.ewald_rc_rule(216, 16.92, 30.0) == pytest.approx(8.46)      # L/2 wins
    assert mdgpu.ewald_rc_rule(32768, 100.64, 12.0) == pytest.approx(12.0)  # cap
    with pytest.raises(mdgpu.MdgError):
        mdgpu.ewald_rc_rule(0, 10.0, 12.0)


def test_no_device_fails_loudly():
    if mdgpu.device_count() > 0:
        pytest.skip("a device is present")
    with pytest.raises(mdgpu.MdgError) as e:
        mdgpu.Batch([mdgpu.box_params(2, 2, 2, 5.64, 0.5, 27, 5.0)])
    assert e.value.code == mdgpu.E_NODEVICE


def _tune_key(p, precision, nk, cells_ok):
    N = 8 * p.nx * p.ny * p.nz
    return (precision, p.real_mode, int(cells_ok), int(np.floor(2 * np.log2(N) + 0.5)),
            int(np.floor(2 * np.log2(nk) + 0.5)))


def test_tuning_table_round_trip(tmp_path):
    """A loaded table entry is what boxes of that class are created with
    (within what the box allows), saves back unchanged, and clears."""
    mdgpu.tune_clear()
    p = mdgpu.box_params(6, 6, 6, 5.64, 0.35, 200, 10.0, 2.0, real_mode=mdgpu.ANALYTIC)
    t0 = mdgpu.box_tiling(p, mdgpu.MIXED)
    assert t0["tuned"] == 0 and t0["rsplit"] >= 1 and t0["nchunk"] >= 1 and t0["ngroup"] >= 1
    nk = mdref.lib().orc_count_kvectors(200)
    L = 6 * 5.64
    cells_ok = int(np.floor(L / (10.0 * (1 + 1e-9)))) >= 3
    key = _tune_key(p, mdgpu.MIXED, nk, cells_ok)
    src = tmp_path / "t.txt"
    src.write_text("# test table\n" + " ".join(map(str, key)) + "  1 1.25 9 4 8\n")
    mdgpu.tune_load(str(src))
    t1 = mdgpu.box_tiling(p, mdgpu.MIXED)
    assert t1["tuned"] == 1 and t1["cells"] == 1 and t1["rsplit"] == 9
    assert t1["nchunk"] == 4 and t1["ngroup"] == 8
    # 33.84 A / (10 A * 1.25) < 3 cells: the width falls back to rc
    assert t1["cell_w"] == 1.0
    # the other precision is another class
    assert mdgpu.box_tiling(p, mdgpu.FP64)["tuned"] == 0
    out = tmp_path / "u.txt"
    mdgpu.tune_save(str(out))
    rows = [ln.split() for ln in out.read_text().splitlines() if not ln.startswith("#")]
    assert rows == [list(map(str, key)) + ["1", "1.25", "9", "4", "8"]]
    mdgpu.tune_clear()
    assert mdgpu.box_tiling(p, mdgpu.MIXED)["tuned"] == 0
    bad = tmp_path / "bad.txt"
    bad.write_text("1 2 3\n")
    with pytest.raises(mdgpu.MdgError):
        mdgpu.tune_load(str(bad))
    assert mdgpu.box_tiling(p, mdgpu.MIXED)["tuned"] == 0


def test_tune_without_device_fails_loudly():
    if mdgpu.device_count() > 0:
        pytest.skip("a device is present")
    p = mdgpu.box_params(2, 2, 2, 5.64, 0.5, 27, 5.0)
    with pytest.raises(mdgpu.MdgError) as e:
        mdgpu.tune([p], mdgpu.MIXED)
    assert e.value.code == mdgpu.E_NODEVICE


def test_empty_batch_is_an_argument_error():
    with pytest.raises(mdgpu.MdgError) as e:
        mdgpu.Batch([])
    assert e.value.code == mdgpu.E_ARG



def test_pr_library_exports_what_its_header_declares():
    """libmdgpu_pr.so (pr-mdc's routines) exports exactly include/prmdc_abi.h."""
    import prmdgpu
    src = re.sub(r"/\*.*?\*/", "", open(os.path.join(INC, "prmdc_abi.h")).read(), flags=re.S)
    decl = set(re.findall(r"^\s*(?:int|void)\s+(\w+)\s*\(", src, flags=re.M))
    decl |= set(re.findall(r"^extern\s+\w+\s+(\w+);", src, flags=re.M))
    assert {"real_space", "ewald1", "ewald2", "ewald3", "calc_alpha", "cell", "pt"} <= decl
    out = subprocess.run(["nm", "-D", "--defined-only", prmdgpu.LIB_PATH], check=True,
                         capture_output=True, text=True).stdout
    exp = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    assert decl == exp, (sorted(decl - exp), sorted(exp - decl))
    prmdgpu.lib()


def _pr_layout(tmp_path, prmdc_headers):
    exe = tmp_path / ("pr_probe_ref" if prmdc_headers else "pr_probe_abi")
    cmd = ["gcc", "-std=c99", "-fcommon", "-o", str(exe),
           os.path.join(ROOT, "tests", "prmdc_layout.c")]
    cmd += ["-DUSE_PRMDC_HEADERS", "-I" + PRMDC_PROGRAM] if prmdc_headers else ["-I" + INC]
    subprocess.run(cmd, check=True, capture_output=True)
    return subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout


def test_pr_ctypes_mirror_matches_header_layout(tmp_path):
    import prmdgpu
    lines = dict(ln.rsplit(" ", 1) for ln in _pr_layout(tmp_path, False).splitlines())
    for cname, cls in (("calc_control", prmdgpu.CalcControl),
                       ("cell_parameters", prmdgpu.CellParameters),
                       ("system_property", prmdgpu.SystemProperty),
                       ("pressure_temperature_set", prmdgpu.PressureTemperature),
                       ("ion_property", prmdgpu.IonProperty),
                       ("Parrinello_Rahman_set", prmdgpu.PrSet)):
        assert int(lines[cname]) == C.sizeof(cls), cname
        for key, off in lines.items():
            if key.startswith(cname + "."):
                assert int(off) == getattr(cls, key.split(".", 1)[1]).offset, key


@pytest.mark.skipif(not os.path.isdir(PRMDC_PROGRAM), reason="reference checkout absent")
def test_pr_header_layout_is_prmdc_layout(tmp_path):
    assert _pr_layout(tmp_path, False) == _pr_layout(tmp_path, True)


def test_pr_no_device_fails_loudly():
    import prmdgpu
    import prref
    if mdgpu.device_count() > 0:
        pytest.skip("a device is present")
    s = prref.pr_system(n=2)
    s.bind(prmdgpu.lib())
    with pytest.raises(prmdgpu.MdgPrError) as e:
        prmdgpu.call("real_space")
    assert e.value.code == prmdgpu.E_NODEVICE
