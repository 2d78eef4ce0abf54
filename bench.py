#!/usr/bin/env python3
"""Benchmark of mdc's force path on one MI355X.

Headline (the batched configuration): ion-steps/s of 256 alkali-halide
boxes (64..4096 ions, NaCl or KCl, 300..2000 K) advanced together in the
mixed-precision batched engine (one hipGraph replay covers 8 steps of all
boxes), next to the plain FP64 path run one box at a time on the same
boxes, the relative force error of the mixed path against FP64, its NVE
energy drift, and the reference's CPU path timed on this host.  Beside it,
one leg per other configuration at the size and length it names:

  nacl216    216-ion NaCl, 1000 FP64 NVE steps: mdc's own program linked
             to libmdgpu (the drop-in) and the handle API, next to mdc's
             all-CPU program on the same input
  nacl4096   the 4096-ion NaCl melt at 1500 K, 5000 FP64 NVE steps with
             cell lists, checked against mdc's CPU forces, its FP64
             roofline, and mdc's CPU step time
  kcl32768   32768-ion KCl at 1300 K, 1e4 mixed NVE steps: step time,
             force error and energy drift against FP64 from the same seed
  stream     the serving stream at its stated job lengths, at 50 jobs/s
             and with every job arriving at once (capacity)

    python bench.py [--gpus 1] [--steps 2000] [--warmup 20]

A step is one velocity-Verlet step (mdc's next_rv_verlet + calc_kin +
calc_press) of every box.  Prints one JSON line.
"""
import argparse
import glob
import hashlib
import json
import multiprocessing as mp
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "molecular-dynamics-in-c_amd")
sys.path.insert(0, PKG)
# the autotuner's table for the 256-box workload's box classes on MI355X
DEFAULT_TUNE = os.path.join(PKG, "tuning", "batch256_mi355x.txt")
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "batch256_mixed_traffic.json")

import numpy as np  # noqa: E402

import mdgpu  # noqa: E402
import workloads as W  # noqa: E402

METRIC = "ion-steps/s batched vs plain (1 GPU, rel. force err≤1e-5); p50/p99 latency"
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
# algorithmic FLOP per unit (add/sub/mul/div = 1, fma = 2; transcendental
# calls counted apart): per (k, ion) 9 in the structure-factor pass (phase 5,
# Z cos / Z sin 2, accumulation 2) and 10 in the back-projection; per
# accepted unordered real-space pair 26 in mdc's table mode.
FLOP_SK = 9
FLOP_KF = 10
FLOP_PAIR = 26
# Analytic real space (the mixed path's erfc + BMH pair function), per
# accepted unordered pair: squared distance 8, r with its residual 7, the
# FP32 erfc 25 and its use 3, Gaussian 3, split Born-Mayer exponent 11,
# r^-6 / r^-8 terms 3, potential 6, force 12, accumulation into both ions
# 12, potential and diagonal virial 7: 97 FLOP (+ rsqrt, rcp, 3 exp2).
FLOP_PAIR_ANALYTIC = 97
DTYPE_MIXED = ("FP32 pair arithmetic; k-space products on FP16 hi/lo MFMA (about 22 bits) "
               "with FP32 sums over ion chunks of at most 512 ions and FP64 across chunks; "
               "FP64 positions, minimum image, cutoff test, per-ion force sums, energies, "
               "virial and self term")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def n_of(p):
    return 8 * p.nx * p.ny * p.nz


def host_info():
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": usable}


def sources_sha256():
    """Digest of the kernel and engine sources (molecular-dynamics-in-c_amd/
    csrc), which fix every kernel's memory traffic: a counter summary is
    current only for the sources it was measured on."""
    h = hashlib.sha256()
    for path in sorted(glob.glob(os.path.join(PKG, "csrc", "*"))):
        if path.endswith((".h", ".hip", ".cpp")):
            h.update(os.path.basename(path).encode())
            with open(path, "rb") as f:
                h.update(f.read())
    return h.hexdigest()


def load_states(boxes, temps):
    return [W.initial_state(p, t, seed=1000 + k) for k, (p, t) in enumerate(zip(boxes, temps))]


def make_batch(boxes, states, precision):
    b = mdgpu.Batch(boxes, precision)
    for k, (r, v, sp) in enumerate(states):
        b.set_state(k, r=r, v=v, species=sp)
    return b


def as_table(p):
    q = mdgpu.BoxParams.from_buffer_copy(p)
    q.real_mode = mdgpu.TABLE
    return q


def force_error(boxes, states_now):
    """Mixed vs FP64 forces (same analytic real-space mode) on the same states."""
    out = {}
    for prec in (mdgpu.MIXED, mdgpu.FP64):
        b = make_batch(boxes, states_now, prec)
        b.forces()
        out[prec] = [b.state(k)["f"] for k in range(len(boxes))]
        b.close()
    mx, rms = 0.0, 0.0
    for f32, f64 in zip(out[mdgpu.MIXED], out[mdgpu.FP64]):
        mx = max(mx, float(np.abs(f32 - f64).max() / np.abs(f64).max()))
        rms = max(rms, float(np.sqrt(((f32 - f64) ** 2).sum() / (f64 ** 2).sum())))
    return mx, rms


def energy_slopes(boxes, states, prec, steps, every):
    """Consistent total energy (kin + real + recip/4 + self) sampled every
    `every` steps over `steps` NVE steps after 20 settling steps; returns
    the least-squares slope per ion scaled to 1e4 steps, per box, and |E|
    per ion at the first sample."""
    n = np.array([n_of(p) for p in boxes], dtype=float)
    b = make_batch(boxes, states, prec)
    b.prepare()
    b.step(20)
    t, es = [], []
    for s in range(0, steps + 1, every):
        b.sync()
        t.append(s)
        es.append([b.scalars(k)["etot"] for k in range(len(boxes))])
        if s < steps:
            b.step(every)
    b.close()
    es = np.array(es)                      # [samples, boxes]
    tt = np.array(t, dtype=float)
    A = np.vstack([tt, np.ones_like(tt)]).T
    coef = np.linalg.lstsq(A, es, rcond=None)[0][0]   # energy per step, per box
    return coef / n * 1e4, np.abs(es[0]) / n


def nve_drift(boxes, states, steps, every=100, stride=8):
    """NVE energy drift of every `stride`-th box at dt = 2 fs, mixed and
    FP64 from the same start.  Medians / maxima over the boxes, absolute
    (mdc's energy unit, 1e-18 J, per ion) and relative to |E| per ion."""
    sel = list(range(0, len(boxes), stride))
    sub = [boxes[k] for k in sel]
    sst = [states[k] for k in sel]
    d32, e_per_ion = energy_slopes(sub, sst, mdgpu.MIXED, steps, every)
    d64, _ = energy_slopes(sub, sst, mdgpu.FP64, steps, every)
    return {"boxes": len(sub), "steps": steps, "dt_fs": 2.0, "sample_every": every,
            "unit": "1e-18 J per ion per 1e4 steps",
            "mixed_median_abs": float(np.median(np.abs(d32))),
            "mixed_max_abs": float(np.abs(d32).max()),
            "fp64_median_abs": float(np.median(np.abs(d64))),
            "fp64_max_abs": float(np.abs(d64).max()),
            "mixed_minus_fp64_max_abs": float(np.abs(d32 - d64).max()),
            "relative_to_abs_energy_per_ion": {
                "mixed_median": float(np.median(np.abs(d32) / e_per_ion)),
                "mixed_max": float((np.abs(d32) / e_per_ion).max()),
                "fp64_median": float(np.median(np.abs(d64) / e_per_ion)),
                "mixed_minus_fp64_max": float((np.abs(d32 - d64) / e_per_ion).max())}}


def plain_one_at_a_time(boxes, states, steps, tune=True):
    """The plain FP64 path (mdc's tabulated real space) run box by box, its
    tilings tuned for a lone box on the chip.  Each box's step graph is
    captured before its timing starts and replayed once (8 steps) as
    warm-up, so no capture falls inside a timed window."""
    plain = [as_table(p) for p in boxes]
    rep = mdgpu.tune(plain, mdgpu.FP64, context_ions=1) if tune else None
    ion_steps, secs = 0.0, 0.0
    for k, q in enumerate(plain):
        r, v, sp = states[k]
        b = mdgpu.Batch([q], mdgpu.FP64)
        b.set_state(0, r=r, v=v, species=sp)
        b.prepare()
        b.step(8)
        b.sync()
        t0 = time.perf_counter()
        b.step(steps)
        b.sync()
        secs += time.perf_counter() - t0
        ion_steps += b.n_ions(0) * steps
        b.close()
    return ion_steps / secs, secs, rep


# ---------------------------------------------------------------- CPU legs
_CPU_TASKS = []   # (nx, ny, nz, a, alpha, hm, rc, dt, r, v) set before the pool forks


def _cpu_step(task):
    """One MD step of mdc's own code (or the oracle) on one box; returns
    (ions, seconds of the step)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import mdref
    nx, ny, nz, a, alpha, hm, rc, dt, r, v = task
    if mdref.ref_available():
        R = mdref.Reference(nx, ny, nz, a, alpha, hm, rc, dt)
        R.set_state(r=r, v=v, f=np.zeros_like(r))
        t0 = time.perf_counter()
        R.run(1, 1)
    else:
        o = mdref.Oracle(nx, ny, nz, a, alpha, hm, rc, dt, real_mode=0)
        o.set_state(r=r, v=v, f=np.zeros_like(r))
        t0 = time.perf_counter()
        o.nve(1)
    return r.shape[1], time.perf_counter() - t0


def _cpu_worker(idx):
    ions, secs = 0, 0.0
    for k in idx:
        n, s = _cpu_step(_CPU_TASKS[k])
        ions += n
        secs += s
    return ions, secs


def cpu_baseline(boxes, states, budget_s, cores):
    """mdc's own CPU code (oracle/_ref, built from the reference sources) or,
    where that library is absent, the oracle restatement: one MD step per
    box over every 4th box in size order.  Single-threaded as mdc is:
    one core in sequence, then one process per core over the same boxes
    (each process steps its share; the figure is the sum of the processes'
    rates, all running at once).  Run before the GPU is initialised, so the
    worker processes fork from a process without HIP state."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import mdref
    global _CPU_TASKS
    kind = "reference" if mdref.ref_available() else "port"
    order = sorted(range(len(boxes)), key=lambda k: n_of(boxes[k]))
    sample = [k for k in order[2::4] if n_of(boxes[k]) <= 32767]   # mdc's short N
    _CPU_TASKS = [(boxes[k].nx, boxes[k].ny, boxes[k].nz, boxes[k].ax, boxes[k].alpha,
                   boxes[k].hm, boxes[k].rc, boxes[k].dt_fs, states[k][0], states[k][1])
                  for k in sample]
    ions, secs, done = 0, 0.0, 0
    for j in range(len(_CPU_TASKS)):
        n, s = _cpu_step(_CPU_TASKS[j])
        ions += n
        secs += s
        done += 1
        if secs > budget_s:
            break
    one = ions / secs
    allc = None
    if cores > 1:
        ctx = mp.get_context("fork")
        shares = [list(range(w, done, cores)) for w in range(cores)]
        shares = [s for s in shares if s]
        with ctx.Pool(len(shares)) as pool:
            res = pool.map(_cpu_worker, shares)
        allc = sum(i / s for i, s in res if s > 0)
    sizes = [n_of(boxes[k]) for k in sample[:done]]
    info = host_info()
    return dict(value=one, unit="ion-steps/s", cores=1, kind=kind,
                all_cores={"value": allc, "processes": min(cores, done)},
                cpu_model=info["cpu_model"], nproc=info["nproc"],
                usable_cpus=info["usable_cpus"],
                sample=(f"1 MD step (next_rv_verlet+calc_kin+calc_press) of each of {done} of "
                        f"the 256 boxes (every 4th by size, N={min(sizes)}..{max(sizes)}), FP64, "
                        f"mdc's tabulated real space over all pairs and direct S(k) with the "
                        f"boxes' alpha/rc/hm; KCl boxes run with mdc's NaCl parameters (same "
                        f"cost); {secs:.1f} s on one core, then the same boxes on "
                        f"{min(cores, done)} processes at once"))


def leg_nacl216(steps=1000):
    """216-ion NaCl (3x3x3 cells, a = 5.64 A, alpha = 5.6/L, rc = L/2,
    hm = 27, dt = 2 fs), Maxwell velocities at 1000 K from a seed, FP64,
    NVE.  (1) mdc's own program: the same driver (oracle/dropin_main.c)
    linked with all of mdc's sources (mdc_cpu) and with libmdgpu in place
    of real.c, rcprcl.c, rv.c, pt.c and msd.c (mdc_gpu), on the same
    input; per-step time = (t(steps) - t(1)) / (steps - 1), which removes
    process start and device set-up; every printed step compared.
    (2) The handle API on the same box, state resident, graph replays."""
    out = {"workload": "NaCl 216 ions, 3x3x3 cells, a=5.64 A, alpha=5.6/L, rc=L/2, hm=27, "
                       "1000 K Maxwell, dt=2 fs, FP64 NVE", "steps": steps}
    ref = os.path.join(ROOT, "oracle", "_ref")
    progs = {k: os.path.join(ref, k) for k in ("mdc_cpu", "mdc_gpu")}
    if all(os.path.exists(p) for p in progs.values()):
        args = ["3", None, "11", "0", "0", "0", "0", "5.64", "1000"]
        rows, tms = {}, {}
        for name, path in progs.items():
            t = {}
            for n in (1, steps):
                a = list(args)
                a[1] = str(n)
                t0 = time.perf_counter()
                p = subprocess.run([path, *a], capture_output=True, text=True, timeout=600)
                t[n] = time.perf_counter() - t0
                if p.returncode != 0:
                    raise RuntimeError(f"{name} failed: {p.stderr[-500:]}")
            tms[name] = (t[steps] - t[1]) / (steps - 1) * 1e3
            rows[name] = np.array([[float(x) for x in ln.split()[1:]]
                                   for ln in p.stdout.splitlines() if ln[:1].isdigit()])
        c, g = rows["mdc_cpu"], rows["mdc_gpu"]

        def diff(k0, k1):
            return max(float(np.abs(g[k0:k1, k] - c[k0:k1, k]).max() / np.abs(c[k0:k1, k]).max())
                       for k in range(c.shape[1]))

        err20, err = diff(0, 20), diff(0, steps)
        out["mdc_program"] = {"mdc_cpu_ms_per_step": tms["mdc_cpu"],
                              "mdc_gpu_dropin_ms_per_step": tms["mdc_gpu"],
                              "dropin_speedup": tms["mdc_cpu"] / tms["mdc_gpu"],
                              "max_rel_diff_pot_kin_vir_pres_first_20_steps": err20,
                              "max_rel_diff_pot_kin_vir_pres_all_steps": err,
                              "diff_note": "the two trajectories start bit-equal and part at "
                                           "the reciprocal sum's rounding (1e-16); molecular "
                                           "dynamics amplifies that over hundreds of steps",
                              "note": "the drop-in copies mdc's arrays host<->device every "
                                      "routine call (mdc's globals stay authoritative)"}
        log(f"  mdc program: cpu {tms['mdc_cpu']:.3f} ms/step, drop-in {tms['mdc_gpu']:.3f} "
            f"ms/step, max rel diff {err20:.1e} (20 steps), {err:.1e} ({steps} steps)")
    p = W.nacl216(real_mode=mdgpu.TABLE)
    r, v, sp = W.initial_state(p, 1000.0, seed=216)
    b = mdgpu.Batch([p], mdgpu.FP64)
    b.set_state(0, r=r, v=v, species=sp)
    b.prepare()
    b.step(8)
    b.sync()
    t0 = time.perf_counter()
    b.step(steps)
    b.sync()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    b.close()
    out["handle_api"] = {"ms_per_step": ms, "ion_steps_per_s": 216 / (ms * 1e-3)}
    log(f"  handle API: {ms:.4f} ms/step")
    return out


def fp64_roofline(prof, work):
    """The FP64 path's kernels against the FP64 peak: algorithmic FLOP per
    launch (per (k, ion) 9 and 10, per accepted pair 26) over the average
    launch duration of profile()."""
    kn = sum(w["nk"] * w["n_ions"] for w in work)
    pairs = sum(w["pairs"] for w in work)
    flops = {"sk_partial": FLOP_SK * kn, "kforce": FLOP_KF * kn, "real_space": FLOP_PAIR * pairs}
    per = {}
    for k, f in flops.items():
        if k not in prof:
            continue
        ms_step, launches = prof[k]
        avg = ms_step / launches
        ach = f / launches / (avg * 1e-3) / 1e12
        per[k] = {"achieved": ach, "frac": ach / PEAK_FP64_TFLOPS, "avg_ms": avg,
                  "launches_per_step": launches, "flop_per_launch": f / launches}
    dom = max(prof, key=lambda k: prof[k][0])
    return dom, per


def leg_nacl4096(steps=5000, cpu_steps=10):
    """The 4096-ion NaCl melt (workloads.melt_snapshot: 3000 K then
    1500 K under mdc's velocity rescaling), then `steps` FP64 NVE steps
    with mdc's table and cell lists (rc = 10 A, Ewald at 1e-6), timed with
    the state resident.  The forces at the snapshot are compared with
    mdc's own calc_foc (oracle/_ref) or the oracle; mdc's CPU step time is
    measured over cpu_steps steps from the same snapshot."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import mdref
    p = W.nacl4096(real_mode=mdgpu.TABLE)
    t0 = time.perf_counter()
    r, v, sp, msd = W.melt_snapshot(p)
    melt_s = time.perf_counter() - t0
    # the autotuner's tilings for this box alone on the chip
    tune = mdgpu.tune([p], mdgpu.FP64, context_ions=1)
    tiling = mdgpu.box_tiling(p, mdgpu.FP64)
    b = mdgpu.Batch([p], mdgpu.FP64)
    b.set_state(0, r=r, v=v, species=sp)
    b.forces()
    f_gpu = b.state(0)["f"]
    b.set_state(0, r=r, v=v)
    b.prepare()
    b.sync()
    e0 = b.scalars(0)
    t0 = time.perf_counter()
    b.step(steps)
    b.sync()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    e1 = b.scalars(0)
    prof = b.profile(10)
    work = [b.work(0)]
    b.close()
    dom, per = fp64_roofline(prof, work)
    kind = "reference" if mdref.ref_available() else "port"
    if kind == "reference":
        R = mdref.Reference(8, 8, 8, p.ax, p.alpha, p.hm, p.rc, p.dt_fs)
        R.set_state(r=r, v=v, f=np.zeros_like(r))
        R.call("calc_foc")
        f_cpu = R.state()["f"]
        R.set_state(r=r, v=v, f=np.zeros_like(r))
        t0 = time.perf_counter()
        R.run(1, cpu_steps)
        cpu_ms = (time.perf_counter() - t0) * 1e3 / cpu_steps
    else:
        o = mdref.Oracle(8, 8, 8, p.ax, p.alpha, p.hm, p.rc, p.dt_fs, real_mode=0)
        o.set_state(r=r, v=v, f=np.zeros_like(r))
        o.calc_foc()
        f_cpu = o.f.copy()
        o.set_state(r=r, v=v, f=np.zeros_like(r))
        t0 = time.perf_counter()
        o.nve(cpu_steps)
        cpu_ms = (time.perf_counter() - t0) * 1e3 / cpu_steps
    err = float(np.abs(f_gpu - f_cpu).max() / np.abs(f_cpu).max())
    log(f"  nacl4096 melt: {ms:.3f} ms/step FP64, CPU {cpu_ms:.0f} ms/step ({kind}), "
        f"force err vs CPU {err:.1e}, MSD {msd[0]:.2f}/{msd[1]:.2f} A^2")
    return {"workload": "NaCl 4096 ions (8x8x8 cells, a=5.64 A), melted at 1500 K "
                        "(3000 steps at 3000 K, 1000 at 1500 K under mdc's rescaling, "
                        "mixed precision), cell lists rc=10 A, Ewald 1e-6 "
                        f"(alpha={p.alpha:.4f}, hm={p.hm}), FP64 table, NVE",
            "melt": {"seconds": melt_s, "msd_last_2ps_A2": [float(x) for x in msd]},
            "autotune": {"seconds": tune["seconds"], "tiling": tiling},
            "steps": steps, "ms_per_step": ms, "ion_steps_per_s": 4096 / (ms * 1e-3),
            "max_rel_force_err_vs_cpu_at_snapshot": err,
            "energy_change_per_ion": (e1["etot"] - e0["etot"]) / 4096,
            "energy_unit": "1e-18 J (table mode: the tabulated force is not the gradient "
                           "of the tabulated potential)",
            "dominant_kernel": dom, "roofline_fp64": per,
            "cpu": {"kind": kind, "ms_per_step": cpu_ms, "steps_timed": cpu_steps, "cores": 1,
                    "speedup_gpu_vs_cpu": cpu_ms / ms}}


def leg_kcl32768(steps=10000, fp64_steps=1000, every=500, cpu=True):
    """32768-ion KCl (16^3 cells, a = 6.29 A) at 1300 K, mixed precision
    (analytic erfc + BMH, cell lists, Ewald at 1e-6 with the cutoff of the
    library's rule, 16.5 A, in place of 10 A: a fifth of the k vectors), `steps`
    NVE steps timed resident; the force error against FP64 on the same
    start; the energy drift over the run, and FP64's over fp64_steps from
    the same seed (shortened: an FP64 step of this box costs several times
    a mixed one).  mdc cannot hold 32768 ions (its N is a short), so the
    CPU time per step is the oracle's, extrapolated: one step timed at two
    small reciprocal cutoffs, and its k-space part scaled linearly to the
    box's k count."""
    p = W.kcl32768_split()
    r, v, sp = W.initial_state(p, 1300.0, seed=33, jitter=0.2)
    # tilings tuned for the box alone on the chip (context of one copy)
    tune = {prec: mdgpu.tune([p], prec, context_ions=1) for prec in (mdgpu.MIXED, mdgpu.FP64)}
    f = {}
    for prec in (mdgpu.MIXED, mdgpu.FP64):
        b = mdgpu.Batch([p], prec)
        b.set_state(0, r=r, v=v, species=sp)
        b.forces()
        f[prec] = b.state(0)["f"]
        b.close()
    d = f[mdgpu.MIXED] - f[mdgpu.FP64]
    emax = float(np.abs(d).max() / np.abs(f[mdgpu.FP64]).max())
    erms = float(np.sqrt((d ** 2).sum() / (f[mdgpu.FP64] ** 2).sum()))
    res = {}
    for prec, n in ((mdgpu.MIXED, steps), (mdgpu.FP64, fp64_steps)):
        b = mdgpu.Batch([p], prec)
        b.set_state(0, r=r, v=v, species=sp)
        b.prepare()
        b.step(20)           # the first energies need a force evaluation
        t, es, secs = [], [], 0.0
        for s in range(0, n + 1, max(1, min(every, n // 2))):   # >= 3 samples for the fit
            b.sync()
            t.append(s)
            es.append(b.scalars(0)["etot"])
            if s < n:
                t0 = time.perf_counter()
                b.step(max(1, min(every, n // 2)))
                b.sync()
                secs += time.perf_counter() - t0
        work = b.work(0)
        b.close()
        tt = np.array(t, dtype=float)
        slope = np.polyfit(tt, np.array(es), 1)[0] / 32768 * 1e4
        res[prec] = {"ms_per_step": secs * 1e3 / n, "steps": n, "drift_abs": float(slope),
                     "drift_rel": float(abs(slope) / (abs(es[0]) / 32768))}
        log(f"  kcl32768 {'mixed' if prec else 'fp64'}: {secs * 1e3 / n:.3f} ms/step, drift "
            f"{slope:.2e} per ion per 1e4 steps")
    out = {"workload": f"KCl 32768 ions (16x16x16 cells, a=6.29 A), 1300 K, Ewald 1e-6 with "
                       f"rc={p.rc:.2f} A from mdg_ewald_rc_rule (alpha={p.alpha:.4f}, hm={p.hm}, "
                       f"{work['nk']} k), cell lists, NVE dt=2 fs",
           "ms_per_step": res[mdgpu.MIXED]["ms_per_step"],
           "ion_steps_per_s": 32768 / (res[mdgpu.MIXED]["ms_per_step"] * 1e-3),
           "steps": steps, "max_rel_force_err_vs_fp64": emax, "rms_rel_force_err_vs_fp64": erms,
           "drift_per_ion_per_1e4_steps": {"unit": "1e-18 J", "mixed": res[mdgpu.MIXED]["drift_abs"],
                                           "fp64": res[mdgpu.FP64]["drift_abs"],
                                           "mixed_rel_to_abs_energy": res[mdgpu.MIXED]["drift_rel"],
                                           "fp64_rel_to_abs_energy": res[mdgpu.FP64]["drift_rel"],
                                           "fp64_steps": fp64_steps,
                                           "fp64_shortened_by": steps / fp64_steps},
           "fp64_ms_per_step": res[mdgpu.FP64]["ms_per_step"],
           "autotune": {"mixed": tune[mdgpu.MIXED], "fp64": tune[mdgpu.FP64]}}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import mdref
        ts = {}
        for hm in (4, 64):
            o = mdref.Oracle(16, 16, 16, p.ax, p.alpha, hm, p.rc, p.dt_fs, salt="KCl", real_mode=0)
            o.set_state(r=r, v=v, f=np.zeros_like(r))
            t0 = time.perf_counter()
            o.nve(1)
            ts[mdref.lib().orc_count_kvectors(hm)] = time.perf_counter() - t0
        (k1, t1), (k2, t2) = sorted(ts.items())
        per_k = (t2 - t1) / (k2 - k1)
        est = t1 + per_k * (work["nk"] - k1)
        out["cpu"] = {"kind": "port (oracle; mdc's short N cannot hold 32768 ions)",
                      "ms_per_step_extrapolated": est * 1e3, "cores": 1,
                      "how": f"one step timed at hm=4 ({k1} k, {t1:.2f} s) and hm=64 ({k2} k, "
                             f"{t2:.2f} s); k-space cost scaled linearly to {work['nk']} k",
                      "speedup_gpu_vs_cpu": est * 1e3 / out["ms_per_step"]}
        log(f"  kcl32768 CPU (extrapolated): {est:.1f} s/step")
    return out


def serving(njobs, steps_scale, budget_gb, tune=True):
    """The serving workload: a seeded Poisson stream (50 jobs/s) of NaCl
    boxes, 216..32768 ions log-uniform, 100..5000 steps (x steps_scale),
    300..2000 K, mixed precision, admitted against an HBM budget with
    continuous batching -- the same stream served one job at a time -- and
    the stream with every job arriving at once, which measures capacity
    rather than the arrival rate."""
    jobs = W.stream(seed=11, njobs=njobs, rate_hz=50.0, n_min=216, n_max=32768,
                    steps_min=100, steps_max=5000, steps_scale=steps_scale)
    out = {}
    res = {}

    def config(mode):
        c = mdgpu.ServeConfig()
        c.hbm_budget = int(budget_gb * (1 << 30))
        c.max_slots = 64
        c.precision = mdgpu.MIXED
        c.real_mode = mdgpu.ANALYTIC
        c.eps, c.rc_max, c.dt_fs = 1e-6, 20.0, 2.0
        c.one_at_a_time = 1 if mode == "one_at_a_time" else 0
        c.realtime = 0 if mode == "saturated" else 1
        c.steps_per_tick = 8
        c.auto_split = 1     # rc by mdg_ewald_rc_rule (10 A .. 16.5 A over 216..32768 ions)
        return c

    if tune:
        prm = [mdgpu.serve_job_params(j, config("batched")) for j in jobs]
        out["autotune"] = mdgpu.tune(prm, mdgpu.MIXED)
        log(f"  tuned {out['autotune']['classes']} job classes in {out['autotune']['seconds']:.1f} s")
    for mode in ("batched", "one_at_a_time", "saturated"):
        c = config(mode)
        r, st = mdgpu.serve(jobs, c)
        res[mode] = r
        out[mode] = {"ion_steps_per_s": st.ion_steps_per_s, "p50_latency_s": st.p50_latency_s,
                     "p99_latency_s": st.p99_latency_s, "mean_latency_s": st.mean_latency_s,
                     "wall_s": st.wall_s, "max_resident": st.max_resident, "ticks": st.ticks,
                     "rejected": st.rejected, "peak_gb": st.peak_bytes / (1 << 30)}
        log(f"  serving {mode}: {st.ion_steps_per_s:.4e} ion-steps/s, p50 {st.p50_latency_s:.3f} s,"
            f" p99 {st.p99_latency_s:.3f} s, wall {st.wall_s:.2f} s")
    same = all(a.final_checksum == b.final_checksum == c.final_checksum
               and a.status == b.status == c.status == mdgpu.OK
               for a, b, c in zip(res["batched"], res["one_at_a_time"], res["saturated"]))
    out["saturated"]["note"] = "every job arrives at t=0: latency is queueing plus service"
    # What bounds the saturated run from below: the jobs' work at the rate
    # the chip steps all of them together as one static batch, and the
    # longest job's steps at its own rate alone on the chip.
    prm = [mdgpu.serve_job_params(j, config("saturated")) for j in jobs]
    work = float(sum(n_of(j) * j.steps for j in jobs))
    mix = mdgpu.Batch(prm, mdgpu.MIXED)
    for k, j in enumerate(jobs):
        r0, v0, sp0 = W.initial_state(prm[k], j.temp_k, seed=j.seed)
        mix.set_state(k, r=r0, v=v0, species=sp0)
    mix.prepare()
    mix.step(16)
    mix.sync()
    t0 = time.perf_counter()
    mix.step(64)
    mix.sync()
    mix_ms = (time.perf_counter() - t0) * 1e3 / 64
    mix.close()
    mix_rate = sum(n_of(j) for j in jobs) / (mix_ms * 1e-3)
    crit = sorted(range(len(jobs)), key=lambda k: -n_of(jobs[k]) * jobs[k].steps)[:3]
    lone_s = 0.0
    for k in crit:
        b = mdgpu.Batch([prm[k]], mdgpu.MIXED)
        r0, v0, sp0 = W.initial_state(prm[k], jobs[k].temp_k, seed=jobs[k].seed)
        b.set_state(0, r=r0, v=v0, species=sp0)
        b.prepare()
        b.step(16)
        b.sync()
        t0 = time.perf_counter()
        b.step(64)
        b.sync()
        lone_s = max(lone_s, (time.perf_counter() - t0) / 64 * jobs[k].steps)
        b.close()
    wall_min = max(work / mix_rate, lone_s)
    out["bounds"] = {
        "static_mix_ion_steps_per_s": mix_rate,
        "static_mix_ms_per_step": mix_ms,
        "longest_job_alone_s": lone_s,
        "wall_lower_bound_s": wall_min,
        "saturated_wall_over_bound": out["saturated"]["wall_s"] / wall_min,
        "saturated_vs_static_mix": out["saturated"]["ion_steps_per_s"] / mix_rate,
        "note": "static mix: every job of the stream resident at once, stepped as one batch; "
                "the bound is the larger of the stream's work at that rate and the longest "
                "job's steps alone"}
    log(f"  serving bounds: static mix {mix_rate:.4e} ion-steps/s ({mix_ms:.3f} ms/step), "
        f"longest job alone {lone_s:.2f} s, wall bound {wall_min:.2f} s")
    out.update({"jobs": njobs, "arrival_rate_hz": 50.0, "ions": "216-32768 log-uniform",
                "ewald_split": "eps 1e-6; rc = 16.5 A (N/32768)^(1/6), at least 10 A, at most L/2 "
                               "(mdg_ewald_rc_rule)",
                "steps": "100-5000" + (f" x {steps_scale}" if steps_scale != 1.0 else ""),
                "hbm_budget_gb": budget_gb, "results_identical_alone_and_in_mix": same,
                "total_ion_steps": float(sum(n_of(j) * j.steps for j in jobs))})
    return out


COPY_PROBE_BYTES = 1 << 30


PR_ROUTINES = ("real_space", "ewald1", "ewald2", "ewald3")


def pr_pairs_within_cutoff(s):
    """Unordered pairs of a pr-mdc system within its cutoff under the
    triclinic minimum image (scaled-coordinate fold)."""
    sc = np.linalg.solve(s.hm.a.T, s.r)
    n, cnt = s.N, 0
    for i0 in range(0, n, 256):
        d = sc[:, i0:i0 + 256, None] - sc[:, None, :]
        d -= np.where(d >= 0.5, 1.0, 0.0)
        d += np.where(d < -0.5, 1.0, 0.0)
        c = np.einsum("ka,kij->aij", s.hm.a, d)
        r2 = (c * c).sum(axis=0)
        idx = np.arange(i0, min(i0 + 256, n))[:, None] < np.arange(n)[None, :]
        cnt += int(((r2 <= s.cutoff * s.cutoff) & idx).sum())
    return cnt


# ewald1's pair pass, FP64 FLOP per ordered pair evaluation (add/sub/mul/div
# = 1; exp and sqrt counted apart): difference, scaled-coordinate minimum
# image and r^2 38 for every evaluation; for one inside the cutoff a further
# 100: Mori's erfc as one fraction 74, the force factor 7, the pair's
# charge product and energy 4, the force 6 and half the virial tensor 9.
PR_FLOP_EVAL = 38
PR_FLOP_ACCEPTED = 100


def pr_roofline(n, npairs, ms):
    """ewald1's routine (its pair kernel and two fixed-order sums) against the
    FP64 peak: every ion's pair from both ions, 2 npairs of them inside."""
    flop = n * (n - 1) * PR_FLOP_EVAL + 2 * npairs * PR_FLOP_ACCEPTED
    achieved = flop / (ms * 1e-3) / 1e12
    return {"kernel": "ewald1 pair pass (k_pr_pairs<1> + sums)", "bound": "fp64",
            "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS, "flop_per_launch": flop}


def prmdc_resident(prref, prmdgpu, n, steps=200, cpu_steps=3):
    """pr-mdc's whole step loop (mdg_pr_run: forces, velocity Verlet,
    Parrinello-Rahman cell, thermostat, tensors) resident on the device,
    against pr-mdc's own step loop compiled from its sources
    (oracle/_ref/libprmdcref.so) on one host core, from the same start
    (pr-mdc's NaCl, n^3 cells sheared as the force leg, seed 17)."""
    if not os.path.exists(prref.PREF_SO):
        return None
    import ctypes as C
    R = prref.ref()
    shear = (0.1, 0.05, -0.08)
    for nm in ("next_rv_verlet", "control_press_PR", "calc_kin", "calc_press"):
        getattr(R, nm).argtypes, getattr(R, nm).restype = [], None
    R.control_temp.argtypes, R.control_temp.restype = [C.c_int, C.c_double], None
    R.prref_setup(n, 5.63, *shear, 17, 0.2, 0.05)
    _, _, sysr, ptr, _ = prmdgpu.globals_of(R)
    t0 = time.perf_counter()
    for k in range(1, cpu_steps + 1):
        sysr.step = k
        R.next_rv_verlet()
        R.control_press_PR()
        R.control_temp(ptr.t_control_step, ptr.temp_set)
        R.calc_kin()
        R.calc_press()
    cpu_ms = (time.perf_counter() - t0) * 1e3 / cpu_steps
    R.prref_setup(n, 5.63, *shear, 17, 0.2, 0.05)
    L = prmdgpu.lib()
    prmdgpu.copy_globals(R, L)
    prmdgpu.run(20)
    t0 = time.perf_counter()
    rec = prmdgpu.run(steps)
    gpu_ms = (time.perf_counter() - t0) * 1e3 / steps
    log(f"    resident NPT N={8 * n ** 3}: {gpu_ms:.3f} ms/step, pr-mdc on 1 core "
        f"{cpu_ms:.1f} ms/step")
    return {"ms_per_step": gpu_ms, "steps": steps, "ion_steps_per_s": 8 * n ** 3 / (gpu_ms * 1e-3),
            "cpu_reference_ms_per_step": cpu_ms, "cpu_steps_timed": cpu_steps, "cpu_cores": 1,
            "cpu_kind": "reference (pr-mdc sources)", "speedup_vs_cpu": cpu_ms / gpu_ms,
            "volume_first_last": [float(rec[0, 3]), float(rec[-1, 3])]}


def prmdc_leg():
    """pr-mdc's force routines (libmdgpu_pr: real_space, ewald1, ewald2,
    ewald3, FP64) on NaCl in a sheared (triclinic) cell at pr-mdc's own
    cutoff (half the smallest face distance), alpha (3.8 / that distance)
    and hm = 27: the device-resident force evaluation, the drop-in per call
    with its host copies, parity against the oracle, and the oracle's own
    time on one host core."""
    import prmdgpu
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import prref
    rows = []
    for n in (3, 7):
        s = prref.pr_system(n=n, shear=(0.1, 0.05, -0.08), seed=n)
        L = prmdgpu.lib()
        s.f[:] = 0.0
        s.bind(L)
        for name in PR_ROUTINES:
            prmdgpu.call(name)
        f_gpu = s.f.copy()
        s.f[:] = 0.0
        t0 = time.perf_counter()
        f_orc, _, _ = prref.oracle_run(s, PR_ROUTINES)
        cpu_ms = (time.perf_counter() - t0) * 1e3
        err = float(np.abs(f_gpu - f_orc).max() / np.abs(f_orc).max())
        reps = 20
        t0 = time.perf_counter()
        for _ in range(reps):
            for name in PR_ROUTINES:
                prmdgpu.call(name)
        dropin_ms = (time.perf_counter() - t0) * 1e3 / reps
        ms, per = prmdgpu.profile(50)
        npairs = pr_pairs_within_cutoff(s)
        rows.append({
            "ions": s.N, "cells": f"{n}x{n}x{n}", "cutoff_A": s.cutoff, "alpha": s.a1,
            "device_ms_per_force_eval": ms,
            "device_ms_per_routine": dict(zip(PR_ROUTINES, per)),
            "dropin_ms_per_force_eval_with_host_copies": dropin_ms,
            "cpu_oracle_ms_per_force_eval": cpu_ms, "cpu_cores": 1,
            "device_speedup_vs_cpu_oracle": cpu_ms / ms,
            "pairs_within_cutoff": npairs,
            "pair_terms_per_s_device": 2 * npairs / ((per[0] + per[1]) * 1e-3),
            "max_rel_force_err_vs_oracle": err,
            "roofline": pr_roofline(s.N, npairs, per[1])})
        log(f"  pr-mdc N={s.N}: device {ms:.3f} ms/eval, drop-in {dropin_ms:.3f} ms, "
            f"oracle {cpu_ms:.1f} ms, err {err:.1e}")
        rows[-1]["resident_npt_step"] = prmdc_resident(prref, prmdgpu, n)
    return {"workload": "pr-mdc NaCl (5.63 A cells), cell sheared b += 0.1 a, "
                        "c += 0.05 a - 0.08 b, pr-mdc's cutoff/alpha, hm = 27, FP64",
            "sizes": rows}


def copy_probe():
    """libmdgpu's plain copy kernel (k_copy_probe) over 1 GiB, three times:
    the kernel of known traffic (1 GiB read, 1 GiB written per dispatch)
    that scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters, and the HBM
    rate a streaming kernel reaches here."""
    ms = mdgpu.copy_probe(COPY_PROBE_BYTES, reps=2)
    return {"bytes_read": COPY_PROBE_BYTES, "bytes_written": COPY_PROBE_BYTES, "ms": ms,
            "GBps": 2 * COPY_PROBE_BYTES / (ms * 1e-3) / 1e9}


def traffic_of(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (counters need runs of their own), or None with the reason when
    that summary was measured on other kernel sources than these."""
    if not os.path.exists(TRAFFIC_JSON):
        return None, "no counter summary"
    t = json.load(open(TRAFFIC_JSON))
    if t.get("sources_sha256") != sources_sha256():
        log("WARNING: profiles/batch256_mixed_traffic.json was measured on other kernel "
            "sources; traffic is reported as null")
        return None, "stale: measured on other kernel sources"
    k = t.get("kernels", {}).get(kernel)
    if not k:
        return None, "kernel not in the counter summary"
    return k["traffic_bytes_per_launch"], t.get("source")


def mixed_roofline(prof, work, prof_alone=None):
    """The batched mixed step's kernels against the FP32 peak: algorithmic
    FLOP per launch over the average launch duration of profile() (every
    launch timed by HIP events its own dispatch stamps, in the step's own
    schedule, where the real space and the two reciprocal chains share the
    chip), and per kernel also over its duration launched alone.  The
    k-space passes compute FP32-grade products on FP16 MFMA (three per
    product) and are priced against the FP32 peak."""
    kn = sum(w["nk"] * w["n_ions"] for w in work)
    pairs = sum(w["pairs"] for w in work)
    flops = {"sk_partial": FLOP_SK * kn, "kforce": FLOP_KF * kn,
             "real_space": FLOP_PAIR_ANALYTIC * pairs}
    per = {}
    for k, f in flops.items():
        if k not in prof:
            continue
        ms_step, launches = prof[k]
        avg = ms_step / launches
        ach = f / launches / (avg * 1e-3) / 1e12
        per[k] = {"achieved": ach, "frac": ach / PEAK_FP32_TFLOPS, "avg_ms": avg,
                  "launches_per_step": launches, "flop_per_launch": f / launches}
        if prof_alone and k in prof_alone:
            ms_a, la = prof_alone[k]
            ach_a = f / (ms_a * 1e-3) / 1e12
            per[k].update({"alone_ms_per_step": ms_a, "alone_achieved": ach_a,
                           "alone_frac": ach_a / PEAK_FP32_TFLOPS})
    dom = max(prof, key=lambda k: prof[k][0])
    pipes = {"real_space": "VALU (FP32 pair function, FP64 distances)",
             "kforce": "MFMA f16x3 (hi/lo) + VALU FP32 epilogue",
             "sk_partial": "MFMA f16x3 (hi/lo) + VALU FP32 operands"}
    traffic, src = traffic_of(dom)
    d = per.get(dom, {})
    return {"kernel": dom, "bound": "mfma" if dom != "real_space" else "valu",
            "pipe": pipes.get(dom, "VALU"),
            "achieved": d.get("achieved"), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": d.get("frac"), "traffic": traffic, "traffic_source": src,
            "avg_ms": d.get("avg_ms"), "launches_per_step": d.get("launches_per_step"),
            "flop_per_launch": d.get("flop_per_launch"),
            "per_kernel": per,
            "kernels_ms": {k: round(v[0], 5) for k, v in prof.items()},
            "kernels_alone_ms": ({k: round(v[0], 5) for k, v in prof_alone.items()}
                                 if prof_alone else None),
            "kernels_launches": {k: v[1] for k, v in prof.items()}}


def traffic_summary(fetch_dir, write_dir, out_json):
    """Turns two rocprofv3 --pmc runs of `bench.py --only-batched` (one with
    FETCH_SIZE, one with WRITE_SIZE; their *counter_collection.csv) into
    the per-kernel traffic summary the bench reads: each counter averaged
    over a kernel's dispatches in KiB, scaled by what the same counter
    reads for the copy kernel of known traffic (k_copy_probe, 1 GiB read
    and 1 GiB written), and stamped with the digest of the sources."""
    import csv

    def load(d, counter):
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            raise SystemExit(f"no counter_collection.csv under {d}")
        per = {}
        for fn in files:
            for row in csv.DictReader(open(fn)):
                if row.get("Counter_Name") != counter:
                    continue
                name = row["Kernel_Name"]
                per.setdefault(name, {}).setdefault(row["Dispatch_Id"], 0.0)
                per[name][row["Dispatch_Id"]] += float(row["Counter_Value"])
        return {k: (sum(v.values()) / len(v), len(v)) for k, v in per.items()}

    def short(name):
        return name.split("(")[0].replace("void ", "").replace("mdg::", "").strip()

    fetch, write = load(fetch_dir, "FETCH_SIZE"), load(write_dir, "WRITE_SIZE")
    probe = [k for k in fetch if "k_copy_probe" in k]
    if not probe:
        raise SystemExit("the copy kernel k_copy_probe is not in the FETCH_SIZE run")
    kib = COPY_PROBE_BYTES / 1024
    fs = kib / fetch[probe[0]][0]
    ws = kib / write[[k for k in write if "k_copy_probe" in k][0]][0]
    # kernel -> the family profile() times it under (the cell build and the
    # box scalars are groups of launches: their bytes add up per family launch)
    names = {"k_real_win": "real_space", "k_mixed_tables": "phase_tables",
             "k_sk_mfma": "sk_partial", "k_sk_finalize": "sk_finalize", "k_kf_pack": "kf_pack",
             "k_kforce_mfma": "kforce", "k_force_vv": "force_vv", "k_box_scalars": "box_scalars",
             "k_scal_part": "box_scalars", "k_cell_bin": "cell_build",
             "k_cell_count": "cell_build", "k_cell_scatter": "cell_build",
             "k_cell_place": "cell_build", "k_observe": "observe", "k_snapshot": "snapshot",
             "k_copy_probe": "copy_probe"}
    out = {"source": "profiles/batch256_mixed_hbm_traffic.md", "sources_sha256": sources_sha256(),
           "fetch_scale": fs, "write_scale": ws, "kernels": {}}
    for k, (fk, nd) in fetch.items():
        base = short(k)
        key = next((v for n, v in names.items() if base.startswith(n)), base)
        wk = write.get(k, (0.0, 0))[0]
        e = out["kernels"].setdefault(key, {"kernels": [], "fetch_kib_per_dispatch": 0.0,
                                            "write_kib_per_dispatch": 0.0, "dispatches": nd,
                                            "read_bytes_per_launch": 0,
                                            "write_bytes_per_launch": 0,
                                            "traffic_bytes_per_launch": 0})
        e["kernels"].append(base)
        e["fetch_kib_per_dispatch"] += fk
        e["write_kib_per_dispatch"] += wk
        e["read_bytes_per_launch"] += int(fk * fs * 1024)
        e["write_bytes_per_launch"] += int(wk * ws * 1024)
        e["traffic_bytes_per_launch"] += int((fk * fs + wk * ws) * 1024)
    with open(out_json, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--drift-steps", type=int, default=10000)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-cores", type=int, default=16,
                    help="processes of the all-core CPU baseline (capped by the usable CPUs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the nacl216, nacl4096 and kcl32768 legs")
    ap.add_argument("--nacl4096-steps", type=int, default=5000)
    ap.add_argument("--nacl4096-cpu-steps", type=int, default=10)
    ap.add_argument("--kcl-steps", type=int, default=10000)
    ap.add_argument("--kcl-fp64-steps", type=int, default=1000)
    ap.add_argument("--no-prmdc", action="store_true", help="skip pr-mdc's force routines")
    ap.add_argument("--only-prmdc", action="store_true",
                    help="pr-mdc's force routines alone (prints their JSON object)")
    ap.add_argument("--no-tune", action="store_true", help="built-in tiling rule only")
    ap.add_argument("--tune-file", default=None,
                    help="tuning table: loaded if it exists, else written after tuning "
                         "(default: the committed MI355X table of this workload)")
    ap.add_argument("--retune", action="store_true",
                    help="tune the batch's box classes afresh instead of loading a table")
    ap.add_argument("--only-batched", action="store_true",
                    help="the batched run alone (for profiler runs)")
    ap.add_argument("--serve-jobs", type=int, default=48)
    ap.add_argument("--serve-steps-scale", type=float, default=1.0)
    ap.add_argument("--serve-budget-gb", type=float, default=16.0)
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="write profiles/batch256_mixed_traffic.json from two rocprofv3 "
                         "--pmc runs (FETCH_SIZE, WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.traffic_from:
        traffic_summary(args.traffic_from[0], args.traffic_from[1], TRAFFIC_JSON)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this engine runs on one GPU; boxes would shard across GPUs with no "
              "communication, but multi-GPU runs are not part of this build.", file=sys.stderr)
        sys.exit(2)

    boxes, temps, salts = W.batch256()
    states = load_states(boxes, temps)          # host-side: no HIP call yet
    n_ions = np.array([n_of(p) for p in boxes])
    total_ions = int(n_ions.sum())
    cpu = None
    if not (args.no_cpu or args.only_batched or args.only_prmdc):
        # before the GPU is touched: the all-core leg forks worker processes
        log("CPU baseline")
        cores = max(1, min(args.cpu_cores, host_info()["usable_cpus"] or 1))
        cpu = cpu_baseline(boxes, states, args.cpu_seconds, cores)
        log(f"  {cpu['value']:.4e} ion-steps/s on 1 core, {cpu['all_cores']['value']} on "
            f"{cpu['all_cores']['processes']} processes ({cpu['kind']}, {cpu['cpu_model']})")

    if mdgpu.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.only_prmdc:
        print(json.dumps({"prmdc_triclinic_force_path": prmdc_leg()}), flush=True)
        return
    probe = copy_probe()
    log(f"copy probe: {probe['GBps']:.0f} GB/s")

    # ---- batched mixed precision: the headline ----
    log(f"batched mixed: {len(boxes)} boxes, {total_ions} ions")
    tune_rep = None
    if not args.no_tune:
        tfile = args.tune_file or DEFAULT_TUNE
        if not args.retune and os.path.exists(tfile):
            # mdg_tune's table for these box classes, measured on an MI355X
            # (the tuner times candidates solo; tables from separate runs can
            # settle a class either way within its 2% margin, which moves the
            # concurrent step by several percent, so the bench loads one)
            mdgpu.tune_load(tfile)
            tune_rep = {"loaded": os.path.relpath(tfile, ROOT)}
            log(f"  tuning table {tune_rep['loaded']}")
        else:
            tune_rep = mdgpu.tune(boxes, mdgpu.MIXED)
            if args.tune_file:
                mdgpu.tune_save(args.tune_file)
        tl = [mdgpu.box_tiling(p, mdgpu.MIXED) for p in boxes]
        tune_rep["boxes_with_cells"] = sum(t["cells"] for t in tl)
        tune_rep["rsplit_hist"] = {str(r): sum(t["rsplit"] == r for t in tl)
                                   for r in sorted({t["rsplit"] for t in tl})}
        if "classes" in tune_rep:
            log(f"  tuned {tune_rep['classes']} classes ({tune_rep['runs']} runs) "
                f"in {tune_rep['seconds']:.1f} s")
    b = make_batch(boxes, states, mdgpu.MIXED)
    # the split of the boxes between the step's two concurrent groups, timed
    # on this batch (every box's state is restored afterwards)
    sched_share = None if args.no_tune else b.tune_schedule(64)
    if sched_share is not None:
        log(f"  step schedule: group 0 share {sched_share:.2f}")
    b.prepare()            # the step graph is captured here, outside every timed step
    b.step(args.warmup)
    b.sync()
    e0 = np.array([b.scalars(k)["etot"] for k in range(len(boxes))])
    t0 = time.perf_counter()
    b.step(args.steps)
    b.sync()
    secs = time.perf_counter() - t0
    value = total_ions * args.steps / secs
    e1 = np.array([b.scalars(k)["etot"] for k in range(len(boxes))])
    drift = (e1 - e0) / n_ions / args.steps * 1e4
    log(f"  {secs * 1e3 / args.steps:.3f} ms/step, {value:.4e} ion-steps/s")
    prof = b.profile(10)
    prof_alone = b.profile(10, alone=True)
    work = [b.work(k) for k in range(len(boxes))]
    states_now = []
    for k in range(len(boxes)):
        st = b.state(k)
        states_now.append((st["r"], st["v"], states[k][2]))
    step_loop = None
    if not args.only_batched:
        # the rest of mdc's step loop (main.c:44-60) armed on every box:
        # calc_msd each step, control_temp every 10 steps to the box's
        # temperature, moment_correction every 1000 steps
        for k in range(len(boxes)):
            b.set_observe(k, msd=True, t_control_step=10, temp_K=temps[k],
                          t_control_until=10000, moment_every=1000)
        b.prepare()
        b.step(args.warmup)
        b.sync()
        t0 = time.perf_counter()
        b.step(args.steps)
        b.sync()
        loop_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        loop_prof = b.profile(10)
        step_loop = {"ms_per_step": loop_ms, "ion_steps_per_s": total_ions / (loop_ms * 1e-3),
                     "observe_ms_per_step": loop_prof.get("observe", (None,))[0],
                     "scalings_box0": b.observed(0)["scalings"],
                     "armed": "calc_msd every step, control_temp every 10 steps (box "
                              "temperature), moment_correction every 1000 steps"}
        log(f"  with mdc's loop armed: {loop_ms:.3f} ms/step "
            f"(observe kernel {step_loop['observe_ms_per_step']:.4f} ms)")
        # every box's r, v, scalars and MSD read back to pinned host memory
        # every 10 steps, overlapped with the stepping (the data behind mdc's
        # files/out, positions and velocities)
        nrec = max(10, args.steps // 10 * 10)
        b.sync()
        t0 = time.perf_counter()
        b.step_recorded(nrec, 10, keep=False)
        rec_ms = (time.perf_counter() - t0) * 1e3 / nrec
        step_loop["recorded_every_10_steps_ms_per_step"] = rec_ms
        step_loop["frame_bytes"] = 48 * total_ions
        log(f"  recorded every 10 steps: {rec_ms:.3f} ms/step "
            f"({48 * total_ions / 1e6:.1f} MB per frame)")
    b.close()
    roof = mixed_roofline(prof, work, prof_alone)

    # ---- accuracy of the mixed path ----
    ferr_max = ferr_rms = drift_fit = None
    if not args.only_batched:
        log("force error mixed vs FP64")
        ferr_max, ferr_rms = force_error(boxes, states_now)
        log("NVE drift, mixed and FP64")
        drift_fit = nve_drift(boxes, states, args.drift_steps)
        log(f"  mixed median {drift_fit['mixed_median_abs']:.3e}, fp64 median "
            f"{drift_fit['fp64_median_abs']:.3e}, max |mixed-fp64| "
            f"{drift_fit['mixed_minus_fp64_max_abs']:.3e} (1e-18 J / ion / 1e4 steps)")

    # ---- plain FP64, one box at a time ----
    plain = None
    if not (args.no_plain or args.only_batched):
        log("plain FP64 one box at a time")
        rate, psecs, prep = plain_one_at_a_time(boxes, states, args.plain_steps,
                                                tune=not args.no_tune)
        plain = {"value": rate, "unit": "ion-steps/s", "steps_per_box": args.plain_steps,
                 "warmup_per_box": "graph captured, then 8 steps (one replay)",
                 "real_space": "mdc table, FP64", "seconds": psecs,
                 "autotune_lone_box": prep}
        log(f"  {rate:.4e} ion-steps/s")

    configs = None
    if not (args.no_configs or args.only_batched):
        configs = {}
        log("nacl216 (FP64, 1000 steps)")
        configs["nacl216"] = leg_nacl216()
        log("nacl4096 melt (FP64, cell lists)")
        configs["nacl4096"] = leg_nacl4096(args.nacl4096_steps, args.nacl4096_cpu_steps)
        log("kcl32768 (mixed)")
        configs["kcl32768"] = leg_kcl32768(args.kcl_steps, args.kcl_fp64_steps,
                                           cpu=not args.no_cpu)

    serve = None
    if not (args.no_serve or args.only_batched):
        log("serving stream")
        serve = serving(args.serve_jobs, args.serve_steps_scale, args.serve_budget_gb,
                        tune=not args.no_tune)

    prmdc = None
    if not (args.no_prmdc or args.only_batched):
        log("pr-mdc force routines (libmdgpu_pr)")
        prmdc = prmdc_leg()

    out = {
        "metric": METRIC, "value": value, "unit": "ion-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": secs * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPE_MIXED, "data": "synthetic",
        "config": {"workload": "256 alkali-halide boxes (NaCl/KCl by seed), 64-4096 ions "
                               "log-uniform in whole cells, 300-2000 K, eps 1e-6 Ewald "
                               f"(rc=min({W.BATCH_RC_MAX:g} A, L/2)), analytic erfc+BMH, all boxes per launch",
                   "boxes": len(boxes), "ions": total_ions, "precision": "mixed",
                   "parallelism": "single GPU, boxes batched per launch",
                   "graph": "8-step hipGraph captured and uploaded before the warm-up"},
        "speedup_vs_plain_one_at_a_time": (value / plain["value"]) if plain else None,
        "plain_fp64_one_at_a_time": plain,
        "max_rel_force_err": ferr_max, "rms_rel_force_err": ferr_rms,
        "nve_drift_per_ion_per_1e4_steps": drift_fit,
        "energy_change_timed_run": {"per_ion_per_1e4_steps_median_abs":
                                    float(np.median(np.abs(drift))),
                                    "max_abs": float(np.abs(drift).max()), "unit": "1e-18 J"},
        "autotune": tune_rep,
        "step_schedule": {"group0_share": sched_share,
                          "picked_by": "mdg_batch_tune_schedule (each split timed twice "
                                       "over 64 steps, state restored)"},
        "mdc_step_loop_armed": step_loop,
        "hbm_copy_probe": probe,
        "configs": configs,
        "serving": serve,
        "prmdc_triclinic_force_path": prmdc,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
