#!/usr/bin/env python3
"""Benchmark of mdc's force path on one MI355X.

Headline: ion-steps/s of 256 alkali-halide boxes (64..4096 ions, NaCl or
KCl, 300..2000 K) advanced together in the mixed-precision batched engine
(one hipGraph replay covers 8 steps of all boxes), next to the plain FP64
path run one box at a time on the same boxes, the relative force error of
the mixed path against FP64, its NVE energy drift, and the reference's CPU
path timed on this host.

    python bench.py [--gpus 1] [--steps 2000] [--warmup 20]

A step is one velocity-Verlet step (mdc's next_rv_verlet + calc_kin +
calc_press) of every box.  Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "molecular-dynamics-in-c_amd")
sys.path.insert(0, PKG)
# the autotuner's table for the 256-box workload's box classes on MI355X
DEFAULT_TUNE = os.path.join(PKG, "tuning", "batch256_mi355x.txt")

import numpy as np  # noqa: E402

import mdgpu  # noqa: E402
import workloads as W  # noqa: E402

METRIC = "ion-steps/s batched vs plain (1 GPU, rel. force err≤1e-5); p50/p99 latency"
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
# algorithmic FLOP per unit (add/sub/mul/div = 1, fma = 2; transcendental
# calls counted apart): per (k, ion) 9 in the structure-factor pass (phase 5,
# Z cos / Z sin 2, accumulation 2) and 10 in the back-projection; per
# accepted unordered real-space pair 26.
FLOP_SK = 9
FLOP_KF = 10
FLOP_PAIR = 26
# Analytic real space (the mixed path's erfc + BMH pair function), per
# accepted unordered pair: squared distance 8, r with its residual 7, the
# FP32 erfc 25 and its use 3, Gaussian 3, split Born-Mayer exponent 11,
# r^-6 / r^-8 terms 3, potential 6, force 12, accumulation into both ions
# 12, potential and diagonal virial 7: 97 FLOP (+ rsqrt, rcp, 3 exp2).
FLOP_PAIR_ANALYTIC = 97


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def load_states(boxes, temps):
    return [W.initial_state(p, t, seed=1000 + k) for k, (p, t) in enumerate(zip(boxes, temps))]


def make_batch(boxes, states, precision):
    b = mdgpu.Batch(boxes, precision)
    for k, (r, v, sp) in enumerate(states):
        b.set_state(k, r=r, v=v, species=sp)
    return b


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


def nve_drift(boxes, states, steps, every=100, stride=8):
    """NVE energy drift of every `stride`-th box: the consistent total energy
    sampled every `every` steps over `steps` steps, least-squares slope per
    ion scaled to 1e4 steps, for the mixed and the FP64 path from the same
    start.  Returns medians / maxima over the boxes, in mdc's energy unit
    (1e-18 J) and relative to |E| per ion."""
    sel = list(range(0, len(boxes), stride))
    sub = [boxes[k] for k in sel]
    sst = [states[k] for k in sel]
    n = np.array([8 * p.nx * p.ny * p.nz for p in sub], dtype=float)
    slopes, e_per_ion = {}, None
    for prec in (mdgpu.MIXED, mdgpu.FP64):
        b = make_batch(sub, sst, prec)
        b.step(20)
        t, es = [], []
        for s in range(0, steps + 1, every):
            b.sync()
            t.append(s)
            es.append([b.scalars(k)["etot"] for k in range(len(sub))])
            if s < steps:
                b.step(every)
        b.close()
        es = np.array(es)                      # [samples, boxes]
        tt = np.array(t, dtype=float)
        A = np.vstack([tt, np.ones_like(tt)]).T
        coef = np.linalg.lstsq(A, es, rcond=None)[0][0]   # energy per step, per box
        slopes[prec] = coef / n * 1e4
        if e_per_ion is None:
            e_per_ion = np.abs(es[0]) / n
    d32, d64 = slopes[mdgpu.MIXED], slopes[mdgpu.FP64]
    return {"boxes": len(sub), "steps": steps, "sample_every": every, "unit": "1e-18 J",
            "mixed_median_abs": float(np.median(np.abs(d32))),
            "mixed_max_abs": float(np.abs(d32).max()),
            "fp64_median_abs": float(np.median(np.abs(d64))),
            "fp64_max_abs": float(np.abs(d64).max()),
            "mixed_minus_fp64_max_abs": float(np.abs(d32 - d64).max()),
            "mixed_rel_median_abs": float(np.median(np.abs(d32) / e_per_ion)),
            "mixed_rel_max_abs": float((np.abs(d32) / e_per_ion).max()),
            "mixed_minus_fp64_rel_max_abs": float((np.abs(d32 - d64) / e_per_ion).max())}


def plain_one_at_a_time(boxes, states, steps, warmup, tune=True):
    """The plain FP64 path (mdc's tabulated real space) run box by box, its
    tilings tuned for a lone box on the chip."""
    plain = []
    for p in boxes:
        q = mdgpu.BoxParams.from_buffer_copy(p)
        q.real_mode = mdgpu.TABLE
        plain.append(q)
    rep = mdgpu.tune(plain, mdgpu.FP64, context_ions=1) if tune else None
    ion_steps, secs = 0.0, 0.0
    for k, q in enumerate(plain):
        r, v, sp = states[k]
        b = mdgpu.Batch([q], mdgpu.FP64)
        b.set_state(0, r=r, v=v, species=sp)
        b.step(warmup)
        b.sync()
        t0 = time.perf_counter()
        b.step(steps)
        b.sync()
        secs += time.perf_counter() - t0
        ion_steps += b.n_ions(0) * steps
        b.close()
    return ion_steps / secs, secs, rep


def cpu_baseline(boxes, states, budget_s):
    """mdc's own CPU code (oracle/_ref, built from the reference sources) or,
    where that library is absent, the oracle restatement: one MD step per
    box over every 4th box in size order, single-threaded."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import mdref
    kind = "reference" if mdref.ref_available() else "port"
    order = sorted(range(len(boxes)), key=lambda k: 8 * boxes[k].nx * boxes[k].ny * boxes[k].nz)
    sample = order[2::4]
    ion_steps, secs, done = 0.0, 0.0, 0
    for k in sample:
        p = boxes[k]
        r, v, sp = states[k]
        if kind == "reference":
            R = mdref.Reference(p.nx, p.ny, p.nz, p.ax, p.alpha, p.hm, p.rc, p.dt_fs)
            if R.N != r.shape[1]:
                continue
            R.set_state(r=r, v=v, f=np.zeros_like(r))
            t0 = time.perf_counter()
            R.run(1, 1)
            secs += time.perf_counter() - t0
        else:
            o = mdref.Oracle(p.nx, p.ny, p.nz, p.ax, p.alpha, p.hm, p.rc, p.dt_fs, real_mode=0)
            o.set_state(r=r, v=v, f=np.zeros_like(r))
            t0 = time.perf_counter()
            o.nve(1)
            secs += time.perf_counter() - t0
        ion_steps += r.shape[1]
        done += 1
        if secs > budget_s:
            break
    sizes = [8 * boxes[k].nx * boxes[k].ny * boxes[k].nz for k in sample[:done]]
    return dict(value=ion_steps / secs if secs > 0 else None, unit="ion-steps/s", cores=1,
                kind=kind,
                sample=(f"1 MD step (next_rv_verlet+calc_kin+calc_press) of {done} of the 256 "
                        f"boxes (every 4th by size, N={min(sizes)}..{max(sizes)}), FP64, "
                        f"single thread, mdc's tabulated real space and direct S(k) with the "
                        f"boxes' alpha/rc/hm; KCl boxes run with mdc's NaCl parameters "
                        f"(same cost); {secs:.1f} s of CPU time"))


def serving(njobs, steps_scale, budget_gb, tune=True):
    """The serving workload: a seeded Poisson stream (50 jobs/s) of NaCl
    boxes, 216..32768 ions log-uniform, 100..5000 steps scaled by
    steps_scale, 300..2000 K, mixed precision, admitted against an HBM
    budget with continuous batching -- and the same stream served one job
    at a time."""
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
        c.eps, c.rc_max, c.dt_fs = 1e-6, 10.0, 2.0
        c.one_at_a_time = 1 if mode == "one_at_a_time" else 0
        c.realtime = 1
        c.steps_per_tick = 8
        return c

    if tune:
        prm = [mdgpu.serve_job_params(j, config("batched")) for j in jobs]
        out["autotune"] = mdgpu.tune(prm, mdgpu.MIXED)
        log(f"  tuned {out['autotune']['classes']} job classes in {out['autotune']['seconds']:.1f} s")
    for mode in ("batched", "one_at_a_time"):
        c = config(mode)
        r, st = mdgpu.serve(jobs, c)
        res[mode] = r
        out[mode] = {"ion_steps_per_s": st.ion_steps_per_s, "p50_latency_s": st.p50_latency_s,
                     "p99_latency_s": st.p99_latency_s, "mean_latency_s": st.mean_latency_s,
                     "wall_s": st.wall_s, "max_resident": st.max_resident, "ticks": st.ticks,
                     "rejected": st.rejected, "peak_gb": st.peak_bytes / (1 << 30)}
        log(f"  serving {mode}: {st.ion_steps_per_s:.4e} ion-steps/s, p50 {st.p50_latency_s:.3f} s,"
            f" p99 {st.p99_latency_s:.3f} s")
    same = all(a.final_checksum == b.final_checksum and a.status == b.status == mdgpu.OK
               for a, b in zip(res["batched"], res["one_at_a_time"]))
    out.update({"jobs": njobs, "arrival_rate_hz": 50.0, "ions": "216-32768 log-uniform",
                "steps": f"100-5000 x {steps_scale}", "hbm_budget_gb": budget_gb,
                "results_identical_alone_and_in_mix": same,
                "total_ion_steps": float(sum(8 * j.nx * j.ny * j.nz * j.steps for j in jobs))})
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--drift-steps", type=int, default=2000)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
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
    ap.add_argument("--serve-steps-scale", type=float, default=0.05)
    ap.add_argument("--serve-budget-gb", type=float, default=16.0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this engine runs on one GPU; boxes would shard across GPUs with no "
              "communication, but multi-GPU runs are not part of this build.", file=sys.stderr)
        sys.exit(2)
    if mdgpu.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.only_prmdc:
        print(json.dumps({"prmdc_triclinic_force_path": prmdc_leg()}), flush=True)
        return
    probe = copy_probe()
    log(f"copy probe: {probe['GBps']:.0f} GB/s")

    boxes, temps, salts = W.batch256()
    states = load_states(boxes, temps)
    n_ions = np.array([8 * p.nx * p.ny * p.nz for p in boxes])
    total_ions = int(n_ions.sum())

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
    work = [b.work(k) for k in range(len(boxes))]
    states_now = []
    for k in range(len(boxes)):
        st = b.state(k)
        states_now.append((st["r"], st["v"], states[k][2]))
    # the rest of mdc's step loop (main.c:44-60) armed on every box:
    # calc_msd each step, control_temp every 10 steps to the box's
    # temperature, moment_correction every 1000 steps
    for k in range(len(boxes)):
        b.set_observe(k, msd=True, t_control_step=10, temp_K=temps[k], t_control_until=10000,
                      moment_every=1000)
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
                 "armed": "calc_msd every step, control_temp every 10 steps (box temperature), "
                          "moment_correction every 1000 steps"}
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
    log(f"  recorded every 10 steps: {rec_ms:.3f} ms/step ({48 * total_ions / 1e6:.1f} MB per frame)")
    b.close()

    kn = sum(w["nk"] * w["n_ions"] for w in work)
    pairs = sum(w["pairs"] for w in work)
    flops = {"sk_partial": FLOP_SK * kn, "kforce": FLOP_KF * kn,
             "real_space": FLOP_PAIR_ANALYTIC * pairs}  # the batch runs analytic real space
    # profile(): per kernel, ms per step over its launches and launches per
    # step, timed with HIP events in the step's own schedule (two box groups
    # on separate streams, kernels overlapping), so each launch's duration
    # is the one a kernel trace of this run reports
    dom = max(prof, key=lambda k: prof[k][0])
    dom_step_ms, dom_launches = prof[dom]
    dom_ms = dom_step_ms / dom_launches
    flop_launch = flops[dom] / dom_launches if dom in flops else None
    ach = flop_launch / (dom_ms * 1e-3) / 1e12 if flop_launch else None
    # HBM bytes per launch of the dominant kernel, from the committed
    # rocprofv3 --pmc summary of this workload (counters need their own runs)
    traffic, traffic_src = None, None
    tfile = os.path.join(ROOT, "profiles", "batch256_mixed_traffic.json")
    if os.path.exists(tfile):
        t = json.load(open(tfile))
        if dom in t.get("kernels", {}):
            traffic, traffic_src = t["kernels"][dom]["traffic_bytes_per_launch"], t["source"]
    pipes = {"real_space": "VALU (FP32 pair function, FP64 distances)",
             "kforce": "MFMA f16x3 (hi/lo) + VALU FP32 epilogue",
             "sk_partial": "MFMA f16x3 (hi/lo) + VALU FP32 operands"}
    roof = {"kernel": dom, "bound": "mfma", "pipe": pipes.get(dom, "VALU"),
            "achieved": ach, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": (ach / PEAK_FP32_TFLOPS) if ach else None, "traffic": traffic,
            "traffic_source": traffic_src,
            "avg_ms": dom_ms, "launches_per_step": dom_launches,
            "flop_per_launch": flop_launch,
            "kernels_ms": {k: round(v[0], 5) for k, v in prof.items()},
            "kernels_launches": {k: v[1] for k, v in prof.items()}}

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
        rate, psecs, prep = plain_one_at_a_time(boxes, states, args.plain_steps, 3,
                                                tune=not args.no_tune)
        plain = {"value": rate, "unit": "ion-steps/s", "steps_per_box": args.plain_steps,
                 "real_space": "mdc table, FP64", "seconds": psecs,
                 "autotune_lone_box": prep}
        log(f"  {rate:.4e} ion-steps/s")

    serve = None
    if not (args.no_serve or args.only_batched):
        log("serving stream")
        serve = serving(args.serve_jobs, args.serve_steps_scale, args.serve_budget_gb,
                        tune=not args.no_tune)

    prmdc = None
    if not (args.no_prmdc or args.only_batched):
        log("pr-mdc force routines (libmdgpu_pr)")
        prmdc = prmdc_leg()

    cpu = None
    if not (args.no_cpu or args.only_batched):
        log("CPU baseline")
        cpu = cpu_baseline(boxes, states, args.cpu_seconds)
        log(f"  {cpu['value']:.4e} ion-steps/s ({cpu['kind']})")

    out = {
        "metric": METRIC, "value": value, "unit": "ion-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": secs * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 pair/k-space arithmetic, fp64 positions/sums", "data": "synthetic",
        "config": {"workload": "256 alkali-halide boxes (NaCl/KCl by seed), 64-4096 ions "
                               "log-uniform in whole cells, 300-2000 K, eps 1e-6 Ewald "
                               "(rc=min(10 A, L/2)), analytic erfc+BMH, all boxes per launch",
                   "boxes": len(boxes), "ions": total_ions, "precision": "mixed",
                   "parallelism": "single GPU, boxes batched per launch"},
        "speedup_vs_plain_one_at_a_time": (value / plain["value"]) if plain else None,
        "plain_fp64_one_at_a_time": plain,
        "max_rel_force_err": ferr_max, "rms_rel_force_err": ferr_rms,
        "nve_drift_per_ion_per_1e4_steps": drift_fit,
        "energy_change_timed_run": {"per_ion_per_1e4_steps_median_abs":
                                    float(np.median(np.abs(drift))),
                                    "max_abs": float(np.abs(drift).max()), "unit": "1e-18 J"},
        "autotune": tune_rep,
        "mdc_step_loop_armed": step_loop,
        "hbm_copy_probe": probe,
        "serving": serve,
        "prmdc_triclinic_force_path": prmdc,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
