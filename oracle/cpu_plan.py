"""CPU baseline runs of the oracle (the FP64 restatement of the reference's
step: push with collisions and move, stable cell sort, deposit, forcing,
FFT2Z-preconditioned GMRES to 1e-10), per workload and thread count, with
the time of each phase per step.  Test infrastructure, not product code.

    python oracle/cpu_plan.py --configs 0,1,2 --threads 1,all > cpu_plan.jsonl

Workloads (ions per cell, grid, perturbation, steps): 0 = 32^3 x 8,
sinusoidal-x 1 %, 100 steps; 1 = 128^3 x 32, uniform 1 %; 2 = 256^3 x 16,
uniform 1 %; the physics of the benchmark definition (30/29 amu, 300 K,
B 5e-5 T along z, E0 0.05 V/m along x, nu_in 1e3 s^-1, nu_e 1e5 s^-1,
dt 1e-5 s).  --steps overrides the step count (default: 100 for
workload 0, 3 otherwise); one line of JSON per (workload, threads) run goes
to stdout, a progress line per step to stderr."""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import orc  # noqa: E402

WORKLOADS = {0: (32, 8, 1, 100), 1: (128, 32, 2, 3), 2: (256, 16, 2, 3)}  # N, ppc, perturbation, steps


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def run(cfg, steps, threads):
    N, ppc, pert, _ = WORKLOADS[cfg]
    orc.set_threads(threads)
    Q, ME, KB = orc.Q, orc.ME, orc.KB
    mi, mn = 30 * 1.6605e-27, 29 * 1.6605e-27
    vT = np.sqrt(2 * KB * 300 / mi) / 3
    vnT = np.sqrt(2 * KB * 300 / mn)
    g = orc.Grid((N,) * 3, p1=(0.05 * N,) * 3)
    K = (0.0, 0.0, Q * 5e-5 / ME / 1e5)
    E0 = (0.05, 0.0, 0.0)
    cG = ME * 1e5 / Q
    t0 = time.perf_counter()
    p = orc.init_particles(g, ppc, 0, (0, 0, 0), (vT, vT, vT), pert=pert, amp=0.01)
    perm, offs = orc.sort_perm(g.nc, orc.cell_keys(g, p))
    p = p.permute(perm)
    mom = orc.deposit_sorted(g, p, offs)
    b = orc.rhs(g, K, E0, 0.0, cG, *mom)
    phi, _, _, _ = orc.solve(g, K, mom[0], b, pc=orc.PC_FFT2Z, rtol=1e-10)
    Nc = orc.collision_count(p.np, 1e3, 1e-5, np.sqrt(3) * vT, (0, 0, 0), (0, 0, 0), vnT, mi, mn)
    S = orc.normaliser(p, [0.0, 0.0, 0.0], vnT)
    setup = time.perf_counter() - t0
    ph = dict.fromkeys(("efield", "push", "sort", "permute", "deposit", "rhs", "solve"), 0.0)
    its = 0
    for it in range(steps):
        P = orc.push_params(g, Q / mi, 1e-5, (0, 0, 5e-5), E0, collide=True, mi=mi, mn=mn,
                            viT=np.sqrt(3) * vT, vnT=vnT, Nc=Nc, seed=0, step=it)
        a = time.perf_counter()
        ex, ey, ez = orc.efield(g, E0, phi)
        b1 = time.perf_counter()
        work = p.copy()
        key, S, _ = orc.push(P, ex, ey, ez, work, S)
        b2 = time.perf_counter()
        perm, offs = orc.sort_perm(g.nc, key)
        b3 = time.perf_counter()
        p = work.permute(perm)
        b4 = time.perf_counter()
        mom = orc.deposit_sorted(g, p, offs)
        b5 = time.perf_counter()
        b = orc.rhs(g, K, E0, 0.0, cG, *mom)
        b6 = time.perf_counter()
        phi, k, conv, _ = orc.solve(g, K, mom[0], b, phi0=phi, pc=orc.PC_FFT2Z, rtol=1e-10)
        b7 = time.perf_counter()
        its += k
        for name, dt in zip(ph, (b1 - a, b2 - b1, b3 - b2, b4 - b3, b5 - b4, b6 - b5, b7 - b6)):
            ph[name] += dt
        print(f"cpu_plan: workload {cfg}, {threads} threads, step {it + 1}/{steps}: {b7 - a:.2f} s", file=sys.stderr,
              flush=True)
    total = sum(ph.values())
    return {
        "workload": cfg, "grid": N, "ions_per_cell": ppc, "ions": int(p.np), "threads": threads, "steps": steps,
        "setup_s": setup, "s_per_step": total / steps,
        "particle_steps_per_s": p.np * steps / total,
        "phase_s_per_step": {k: v / steps for k, v in ph.items()},
        "solve_iterations_per_step": its / steps,
        "cpu_model": cpu_model(), "host_threads": os.cpu_count(),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--configs", default="0")
    ap.add_argument("--threads", default="1,all", help="comma list of thread counts; 'all' = every host thread")
    ap.add_argument("--steps", type=int, default=0)
    a = ap.parse_args()
    allt = orc.get_threads()
    for cfg in (int(c) for c in a.configs.split(",")):
        for t in a.threads.split(","):
            n = allt if t == "all" else int(t)
            steps = a.steps or WORKLOADS[cfg][3]
            print(json.dumps(run(cfg, steps, n)), flush=True)


if __name__ == "__main__":
    main()
