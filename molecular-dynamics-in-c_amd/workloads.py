"""Synthetic, seeded inputs for the five workloads of mdc's force path.

  nacl216      NaCl, 216 ions (3x3x3 cells, a = 5.64 A), alpha = 5.6/L,
               cutoff L/2, hm = 27, 1000 K, dt = 2 fs, FP64
  nacl4096     NaCl, 4096 ions (8x8x8), melted at 1500 K (melt_snapshot),
               cutoff 10 A, alpha/hm for a 1e-6 Ewald error, FP64, cell lists
  kcl32768     KCl, 32768 ions (16x16x16, a = 6.29 A), 1300 K, mixed
  batch256     256 alkali-halide boxes (NaCl or KCl by seed), 64..4096 ions
               log-uniform in whole conventional cells, 300..2000 K, mixed
  stream       Poisson stream (mean 50 jobs/s) of NaCl boxes, 216..32768
               ions log-uniform, 100..5000 steps, 300..2000 K

Ewald split for the accuracy-targeted boxes (everything but nacl216): real
cutoff rc = min(rc_max, L_min/2), alpha = s/rc, hm = floor((s alpha L_max/pi)^2),
s = sqrt(-ln eps).  Velocities are Maxwell-Boltzmann from the library's
seeded generator with mdc's centre-of-mass correction.
"""
import math

import numpy as np

import mdgpu

A_NACL = 5.64
A_KCL = 6.29


def ewald_box(nx, ny, nz, a, salt, eps=1e-6, rc_max=10.0, dt_fs=2.0, real_mode=None,
              cells=mdgpu.CELLS_AUTO):
    L = np.array([nx, ny, nz], dtype=float) * a
    rc = min(rc_max, 0.5 * L.min())
    alpha, hm = mdgpu.ewald_params(rc, L.max(), eps)
    if real_mode is None:
        real_mode = mdgpu.ANALYTIC
    return mdgpu.box_params(nx, ny, nz, a, alpha, hm, rc, dt_fs, real_mode=real_mode,
                            cells=cells, salt=salt)


def nacl216(real_mode=mdgpu.TABLE):
    n, a = 3, A_NACL
    L = n * a
    return mdgpu.box_params(n, n, n, a, 5.6 / L, 27, L / 2, 2.0, real_mode=real_mode)


def nacl4096(real_mode=mdgpu.TABLE):
    return ewald_box(8, 8, 8, A_NACL, "NaCl", real_mode=real_mode, cells=mdgpu.CELL_LIST)


def kcl32768(rc_max=10.0):
    return ewald_box(16, 16, 16, A_KCL, "KCl", rc_max=rc_max, cells=mdgpu.CELL_LIST)


def kcl32768_split():
    """kcl32768 with the cutoff of the library's rule (mdg_ewald_rc_rule,
    16.5 A at 32768 ions) instead of 10 A, same 1e-6 Ewald accuracy."""
    return kcl32768(rc_max=mdgpu.ewald_rc_rule(32768, 16 * A_KCL, 30.0))


def cells_for(n_ions):
    """Near-cubic nx x ny x nz conventional cells holding about n_ions."""
    target = max(1.0, n_ions / 8.0)
    c = max(1, int(round(target ** (1.0 / 3.0))))
    nz = max(1, int(round(target / (c * c))))
    return c, c, nz


# Real-space cutoff cap of the batched workload: rc = min(14 A, L/2) at the
# same Ewald accuracy (eps 1e-6).  Measured on the 256 boxes (each cap with
# its own tuning table): 0.924 ms per step at 10 A, 0.829 at 12, 0.780 at
# 14, 0.806 at 17 and 0.965 at L/2; the N^(1/6) rule 22 A (N/32768)^(1/6)
# gives the same as 14 (0.772).  Most of these boxes run all pairs, whose
# screening costs the same at any cutoff, so a longer cutoff buys fewer
# k vectors cheaply.
BATCH_RC_MAX = 14.0


def batch256(seed=2024, nbox=256, eps=1e-6, rc_max=BATCH_RC_MAX):
    """Box parameters and temperatures of the batched workload."""
    rng = np.random.default_rng(seed)
    boxes, temps, salts = [], [], []
    for _ in range(nbox):
        n = math.exp(rng.uniform(math.log(64), math.log(4096)))
        salt = "NaCl" if rng.random() < 0.5 else "KCl"
        a = A_NACL if salt == "NaCl" else A_KCL
        nx, ny, nz = cells_for(n)
        boxes.append(ewald_box(nx, ny, nz, a, salt, eps=eps, rc_max=rc_max))
        temps.append(float(rng.uniform(300.0, 2000.0)))
        salts.append(salt)
    return boxes, temps, salts


def initial_state(p, temp_k, seed, jitter=0.0):
    """mdc's rock-salt lattice (optionally jittered) and seeded velocities."""
    r, sp = mdgpu.rocksalt(p.nx, p.ny, p.nz, (p.ax, p.ay, p.az))
    if jitter > 0:
        L = np.array([p.nx * p.ax, p.ny * p.ay, p.nz * p.az])[:, None]
        rng = np.random.default_rng(seed + 7919)
        r = np.mod(r + rng.uniform(-jitter, jitter, r.shape), L)
    v = mdgpu.maxwell(seed, temp_k, sp, p.pot)
    return r, v, sp


def melt_snapshot(p, seed=42, t_hot=3000.0, hot_steps=3000, t_melt=1500.0, settle_steps=1000):
    """A melt of box p at t_melt, prepared on the device: mdc's lattice
    with Maxwell velocities at t_hot, stepped hot_steps steps (mixed
    precision, analytic erfc + BMH, the box's own Ewald split) under mdc's
    velocity rescaling (control_temp every 10 steps, pt.c:85-112) at t_hot,
    well above the crystal's superheating limit at this density, then
    settle_steps steps rescaled to t_melt.  Returns (r, v, species, msd):
    the state after the last step and the mean-squared displacement per
    species over the settle stage (about 2 ps at dt = 2 fs; a liquid's
    grows to several A^2 there, a crystal's stays below about 0.5 A^2)."""
    q = mdgpu.BoxParams.from_buffer_copy(p)
    q.real_mode = mdgpu.ANALYTIC
    r, v, sp = initial_state(q, t_hot, seed)
    b = mdgpu.Batch([q], mdgpu.MIXED)
    try:
        b.set_state(0, r=r, v=v, species=sp)
        b.set_observe(0, t_control_step=10, temp_K=t_hot, t_control_until=hot_steps)
        b.step(hot_steps)
        b.set_observe(0, msd=True, t_control_step=10, temp_K=t_melt,
                      t_control_until=hot_steps + settle_steps + 1)
        b.step(settle_steps)
        st = b.state(0)
        msd = b.observed(0)["msd"][:2]
    finally:
        b.close()
    return st["r"], st["v"], sp, msd


def stream(seed=7, njobs=64, rate_hz=50.0, n_min=216, n_max=32768, steps_min=100,
           steps_max=5000, steps_scale=1.0):
    """Poisson job stream of NaCl boxes.  steps_scale shortens every job by
    the same factor (for bounded benchmark runs)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for k in range(njobs):
        t += rng.exponential(1.0 / rate_hz)
        n = math.exp(rng.uniform(math.log(n_min), math.log(n_max)))
        nx, ny, nz = cells_for(n)
        j = mdgpu.Job()
        j.arrival_s = t
        j.nx, j.ny, j.nz = nx, ny, nz
        j.temp_k = float(rng.uniform(300.0, 2000.0))
        j.steps = max(1, int(round(rng.integers(steps_min, steps_max + 1) * steps_scale)))
        j.seed = seed * 100003 + k
        jobs.append(j)
    return jobs
