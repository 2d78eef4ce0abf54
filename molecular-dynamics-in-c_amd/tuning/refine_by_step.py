"""Refine a tuning table by the step time of a whole batch.

mdg_tune times each family of kernels solo, on copies of one box class.
Under the concurrent step schedule (two box groups, real space beside the
reciprocal chains) the best tiling can differ.  This tool starts from a
table and does coordinate descent on the 256-box batch's step time: for
each class row it tries halving and doubling the real-space split, the
back-projection groups and the structure-factor chunks, and keeps a change
only if two runs of it beat a fresh run of the current table by 0.7%.

usage: python3 refine_by_step.py [in_table] [out_table]   (GPU required)
"""
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import mdgpu  # noqa: E402
import workloads as W  # noqa: E402

LIMITS = {7: 27, 8: 64, 9: 64}   # columns: rsplit, nchunk, ngroup


def read_table(path):
    hdr, rows = [], []
    for ln in open(path):
        if ln.startswith("#"):
            hdr.append(ln)
            continue
        a = ln.split()
        rows.append([int(x) for x in a[:6]] + [float(a[6])] + [int(x) for x in a[7:10]])
    return hdr, rows


def write_table(hdr, rows, path):
    with open(path, "w") as f:
        f.writelines(hdr)
        for r in rows:
            f.write("%d %d %d %d %d  %d %g %d %d %d\n" % tuple(r))


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "batch256_mi355x.txt")
    dst = sys.argv[2] if len(sys.argv) > 2 else src + ".refined"
    hdr, rows = read_table(src)
    boxes, temps, _ = W.batch256()
    states = [W.initial_state(p, t, seed=1000 + k) for k, (p, t) in enumerate(zip(boxes, temps))]
    scratch = dst + ".cand"

    def measure(rs, steps=200):
        write_table(hdr, rs, scratch)
        mdgpu.tune_clear()
        mdgpu.tune_load(scratch)
        b = mdgpu.Batch(boxes, mdgpu.MIXED)
        for k, (r, v, sp) in enumerate(states):
            b.set_state(k, r=r, v=v, species=sp)
        b.step(16)
        b.sync()
        t0 = time.perf_counter()
        b.step(steps)
        b.sync()
        b.close()
        return (time.perf_counter() - t0) * 1e3 / steps

    best = [list(r) for r in rows]
    cur = min(measure(best), measure(best))
    print(f"start {cur:.4f} ms per step", flush=True)
    for _ in range(2):
        for i in range(len(best)):
            for col in (7, 9, 8):
                for f in (2.0, 0.5):
                    v = best[i][col]
                    nv = max(1, min(LIMITS[col], int(round(v * f))))
                    if nv == v:
                        continue
                    cand = [list(r) for r in best]
                    cand[i][col] = nv
                    ms = measure(cand)
                    if ms < cur * 0.993:
                        ms2, ref = measure(cand), measure(best)
                        if max(ms, ms2) < ref * 0.993:
                            best, cur = cand, min(ms, ms2)
                            print(f"row {i} column {col}: {v} -> {nv}, {cur:.4f} ms", flush=True)
                        else:
                            cur = min(cur, ref)
    os.remove(scratch)
    write_table(hdr, best, dst)
    print(f"final {cur:.4f} ms per step -> {dst}", flush=True)


if __name__ == "__main__":
    main()
