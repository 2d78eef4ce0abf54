"""Writes tests/golden/mdc_nacl216.npz from the compiled mdc sources.

Run in a checkout that has the reference (it needs oracle/_ref/libmdcref.so,
built by `make -C oracle`).  The fixture holds only data: inputs and the
reference's outputs for the NaCl 216-ion box (3x3x3 cells, a = 5.64 A,
alpha = 5.6/L, hm = 27, cutoff L/2, dt = 2 fs):
  * three seeded states (lattice jitter + 1000 K velocities) and, for each,
    the forces, pot and diagonal virial of mdc's calc_foc;
  * a 40-step NVE trajectory from a fourth state (mdc's next_rv_verlet,
    calc_kin, calc_press per step).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import mdref  # noqa: E402


def main():
    a, n = 5.64, 3
    L = n * a
    alpha, hm, rc, dt = 5.6 / L, 27, L / 2, 2.0
    R = mdref.Reference(n, n, n, a, alpha, hm, rc, dt)
    o = mdref.Oracle(n, n, n, a, alpha, hm, rc, dt)   # for masses / lattice only
    rs, vs, fs, sc = [], [], [], []
    for seed in (101, 102, 103):
        r = mdref.jitter_positions(o.r, o.L, 0.25, seed)
        v = mdref.maxwell_velocities(o.masses(), 1000.0, seed)
        R.set_state(r=r, v=v, f=np.zeros_like(r))
        R.call("calc_foc")
        s = R.scalars()
        rs.append(r)
        vs.append(v)
        fs.append(R.state()["f"].copy())
        sc.append([s["pot"], s["virX"], s["virY"], s["virZ"]])
    r0 = mdref.jitter_positions(o.r, o.L, 0.1, 104)
    v0 = mdref.maxwell_velocities(o.masses(), 1000.0, 104)
    steps = 40
    R.set_state(r=r0, v=v0, f=np.zeros_like(r0))
    R.run(1, steps)
    st = R.state()
    np.savez_compressed(
        os.path.join(HERE, "mdc_nacl216.npz"),
        params=np.array([n, a, alpha, hm, rc, dt]), r=np.array(rs), v=np.array(vs),
        f=np.array(fs), scal=np.array(sc), traj_r0=r0, traj_v0=v0, traj_steps=steps,
        traj_r=st["r"], traj_v=st["v"], traj_f=st["f"],
        traj_scal=np.array([R.scalars()[k] for k in ("pot", "virX", "virY", "virZ", "kin")]))
    print("wrote mdc_nacl216.npz")


if __name__ == "__main__":
    main()
