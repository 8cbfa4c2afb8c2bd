"""Time the FP64 oracle's one-box step on the reference's own benchmark input
(TEST INFRASTRUCTURE: the CPU baseline's calibration against the reference
binary).

The reference's main loop (main.cpp:84-139) times 20 steps after 10 warm-up
steps of CIC deposit, transposes, FFT + Green + IFFT and update_particles on
Ng^3 particles drawn by init_particles (utils.cpp:22-36).  This script runs
the oracle's restatement of that step -- deposit, FFT solve with the
reference's Green's function at a = Start_a, centred difference,
trilinear kick (plus the drift the oracle adds) -- on the same particles,
one warm-up step and `--steps` timed steps, with `--threads` OpenMP threads,
and prints seconds per step, so that it can be set beside the reference
binary's seconds per step measured on the same cores.

    python oracle/calibrate.py --ng 512 --threads 8 --steps 3
"""
from __future__ import annotations

import argparse
import json
import time

import pyoracle


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ng", type=int, default=256)
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--steps", type=int, default=3)
    args = ap.parse_args()
    ng = args.ng
    npart = ng ** 3
    pyoracle.set_threads(args.threads)
    pos = list(pyoracle.init_particles(npart, ng))
    vel = [p * 0.0 for p in pos]
    kick = pyoracle.function_a(0.9) * 0.001
    mass = ng ** 3 / npart

    def step():
        delta = pyoracle.deposit_periodic(pos, ng, mass)
        phi = pyoracle.solve(delta, pyoracle.GREEN_REF_KH, 1.0 / 0.9)
        pyoracle.interp_kick_drift(pos, vel, phi, kick, 0.0)

    step()  # warm-up: allocates the private grids and gradient meshes
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    t = (time.perf_counter() - t0) / args.steps
    model = ""
    for line in open("/proc/cpuinfo"):
        if line.startswith("model name"):
            model = line.split(":", 1)[1].strip()
            break
    print(json.dumps({"ng": ng, "np": npart, "threads": args.threads, "steps": args.steps,
                      "oracle_s_per_step": round(t, 3), "cpu": model}))


if __name__ == "__main__":
    main()
