"""Benchmark of the one-box PM step on one MI355X.

Workload (one GPU's share of the benchmark box): 512^3 particles on a 512^3
mesh, Zel'dovich lattice ICs with a k^-2 density spectrum scaled to 1.5-cell
rms displacement (clustered: shell crossing stresses the deposit), growing-
mode velocities, continuous -1/k^2 Green's function (unit box, G = M = 1),
KDK steps of dt = 0.005.  A step = one force evaluation (zero + CIC deposit,
5-pass R2C/Green/C2R FFT solve, interpolation of the centred-difference
force) + kick + drift, all on the GPU with inputs resident in HBM.

Prints one JSON line: value = FP32-mesh Mparticle-steps/s with the
split-precision particle state (positions as integer cell + 2^-23-cell
fixed-point fraction, FP32 velocities, resident on the GPU: pm_split_step);
the FP32 mesh with FP64 positions and velocities (pm_step), the FP64-mesh
path, per-phase times, the roofline of the dominant kernel and the CPU
baseline (the FP64 oracle on the host cores, on a bounded sample of the same
ICs) are reported beside it.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "particle-mesh_amd"))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--ng", type=int, default=512)
    ap.add_argument("--nrow", type=int, default=512)
    ap.add_argument("--rms", type=float, default=1.5)
    ap.add_argument("--dt", type=float, default=0.005)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--deposit", choices=["sorted", "atomic"], default="sorted")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-mesh leg")
    ap.add_argument("--no-fp64-positions", action="store_true",
                    help="skip the FP32-mesh leg with FP64 positions (the headline runs the split state)")
    ap.add_argument("--no-alt-deposit", action="store_true", help="skip the leg with the other deposit variant")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the GPU side of the accuracy leg")
    ap.add_argument("--accuracy-steps", type=int, default=20,
                    help="KDK steps of the accuracy leg and of the timed CPU baseline")
    ap.add_argument("--no-large-adjoint", action="store_true", help="skip the 1024^3-mesh adjoint leg")
    ap.add_argument("--large-nrow", type=int, default=512)
    ap.add_argument("--large-ng", type=int, default=1024)
    ap.add_argument("--large-steps", type=int, default=30)
    ap.add_argument("--no-tune", action="store_true", help="use the default FFT pass shapes (no pm_plan_tune)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = nproc (honours OMP_NUM_THREADS)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 1 GiB device copy (PMC byte-counter calibration)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(ng: int, npart: int, prec: int, staged: bool = True, split: bool = False):
    """Compulsory HBM bytes per launch of each phase of a steady-state step
    (b = mesh real bytes; positions and velocities FP64, 48 B per particle,
    or the split state, 24 B: u32 cell + fraction and FP32 velocity per
    axis; spectrum (ng/2+1) ng^2 complex).  With the sorted deposit the
    interpolation stages every drifted position in tile order as a record of
    its in-tile offsets (12 B: 3 x u32 for an FP32 mesh; 16 B: 3 x 42 bits
    for an FP64 mesh) and the next deposit reads it back: written once and
    read once.  The split state is kept in tile order instead: no record."""
    b = 4 if prec == 0 else 8
    st = (12 if prec == 0 else 16) if staged else 0
    # the split state: 3 u32 + 3 FP32 with an FP32 mesh, 3 u64 + 3 FP64 with an FP64 mesh
    state = (24 if prec == 0 else 48) if split else 48
    nc = ng ** 3
    nk = (ng // 2 + 1) * ng * ng
    if split:
        # the split state lives in tile order: the deposit reads the 12 B of
        # positions from the tile buckets, the interpolation reads and writes
        # the 24 B of state in place (no record)
        st = 0
        dep = 12 if prec == 0 else 24
    else:
        dep = st if staged else 24
    return {
        "deposit": dep * npart + b * nc,                     # read positions; write the mesh once
        "fft_z_fwd": b * nc + 2 * b * nk,                    # real in, half spectrum out
        "fft_y_fwd": 4 * b * nk,                             # spectrum in + out
        "fft_x_green": 4 * b * nk,
        "fft_y_inv": 4 * b * nk,
        "fft_z_inv": 2 * b * nk + b * nc,
        "interp_kick_drift": b * nc + (2 * state + st) * npart,  # read phi; read+write x, v; write staging
    }


# kernel that carries each phase of the FP32 sorted step, as rocprofv3 names it;
# for the FFT passes a name prefix: the lines / threads template arguments that
# follow it are the shape that pm_plan_tune chose
PHASE_KERNELS = {
    "deposit": "void pm::k_deposit_staged<float, 16>",
    "fft_z_fwd": "void pm::k_zfwd<float, 512, true,",
    "fft_y_fwd": "void pm::k_yfwd<float, 512,",
    "fft_x_green": "void pm::k_x<float, 512, 7,",
    "fft_y_inv": "void pm::k_yinv<float, 512,",
    "fft_z_inv": "void pm::k_zinv<float, 512,",
    "interp_kick_drift": "void pm::k_interp<float, 2, 512, 1>",
}
# the same for the split-state step (pm_split_step)
SPLIT_PHASE_KERNELS = dict(PHASE_KERNELS, deposit="void pm::k_deposit_state<16>",
                           interp_kick_drift="void pm::k_state_interp<16, true>")
FP64_SPLIT_INTERP_KERNEL = "void pm::k_state_interp64<16, true>"


def measured_traffic(kernel: str, ng: int, npart: int, table: str = ""):
    """HBM bytes per dispatch of a kernel from the committed rocprofv3 PMC
    summary of this workload (FETCH_SIZE and WRITE_SIZE, each scaled by what
    the counter reads for a plain copy of known size); None if absent."""
    path = ROOT / "profiles" / f"pmc_traffic_{ng}{table}.txt"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("#") or "|" not in line:
            continue
        cols = [c.strip() for c in line.split("|")]
        if len(cols) >= 6 and cols[0] == kernel and cols[1] == str(npart):
            return int(float(cols[5]))
    return None


# kernels of the FP32 deposit phase of each variant and the PMC summary that
# holds their HBM bytes (the sorted step's steady state: staged records)
DEPOSIT_KERNELS = {
    "sorted": ("", ["void pm::k_deposit_staged<float, 16>", "void pm::k_fold_faces<float, 16, 1, 0>",
                    "void pm::k_fold_seams<float, 16, 1, 0>"]),
    "atomic": ("_atomic", ["void pm::k_fill<float>", "void pm::k_deposit_atomic<float>"]),
    "split": ("", ["void pm::k_state_arrive<16>", "void pm::k_deposit_state<16>",
                   "void pm::k_fold_faces<float, 16, 1, 0>", "void pm::k_fold_seams<float, 16, 1, 0>"]),
}


def deposit_traffic(name: str, ng: int, npart: int, phase_ms: float):
    """HBM bytes of the FP32 deposit phase's kernels (committed PMC summary)
    and the phase's HBM rate at the measured phase time; None if absent."""
    table, kernels = DEPOSIT_KERNELS[name]
    per = {k: measured_traffic(k, ng, npart, table) for k in kernels}
    if not phase_ms or any(v is None for v in per.values()):
        return None
    total = sum(per.values())
    return {"kernels_hbm_bytes": per, "hbm_bytes": total, "phase_ms": round(phase_ms, 4),
            "hbm_GBs": round(total / (phase_ms * 1e-3) / 1e9, 1),
            "source": f"profiles/pmc_traffic_{ng}{table}.txt"}


# the FP64 step's dominant kernel, and the PMC summary that holds its traffic
# (columns: kernel | dispatches | FETCH_KiB | WRITE_KiB | hbm_bytes | ...)
FP64_INTERP_KERNEL = "void pm::k_interp<double, 2, 256, 1>"


def measured_traffic_fp64(kernel: str, ng: int):
    path = ROOT / "profiles" / f"pmc_traffic_{ng}_fp64_adjoint.txt"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("#") or "|" not in line:
            continue
        cols = [c.strip() for c in line.split("|")]
        if len(cols) >= 5 and cols[0] == kernel:
            return int(float(cols[4]))
    return None


def make_ics(nrow, ng, rms, seed, dev):
    import torch
    import pmhip

    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    noise = torch.randn(nrow ** 3, dtype=torch.float64, device=dev, generator=g)
    vel_factor = math.sqrt(4.0 * math.pi)  # growing mode of a static background, G = rho_bar = 1
    pos, vel = pmhip.zeldovich_ic(nrow, ng, noise, rms, vel_factor)
    del noise
    torch.cuda.synchronize()
    return pos, vel


def run_gpu(prec, pos0, vel0, args, ng, npart, variant, split=False):
    """Warm up, then time exactly args.steps steps; returns (ms_per_step,
    phase ms per launch).  split: the plan's split-precision state
    (pm_split_load before the warm-up, pm_split_step, pm_split_store after
    the timed steps into the returned arrays)."""
    import torch
    import pmhip

    pos = [p.clone() for p in pos0]
    vel = [v.clone() for v in vel0]
    plan = pmhip.Plan(ng, npart, prec, pmhip.PM_GREEN_CONT_K2, 1.0)
    if not args.no_tune:
        # the FFT pass shapes are chosen by timing them on this plan (and cached
        # for the later plans of this process), as FFTW plans before it runs
        plan.tune(5)
    shape = plan.fft_shape
    dt = args.dt
    step = 0
    if split:
        plan.split_load(pos, vel)

    def one():
        nonlocal step
        kick = 0.5 * dt if step == 0 else dt
        if split:
            plan.split_step(kick, dt)
        else:
            # after the first step the positions are the ones this plan wrote:
            # let the deposit bin from the keys the interpolation left
            plan.step(pos, vel, kick, dt, variant | (pmhip.PM_STEP_REUSE_BINS if step > 0 else 0))
        step += 1

    for _ in range(args.warmup):
        one()
    torch.cuda.synchronize()
    plan.reset_timing()
    plan.enable_timing(args.steps)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        one()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = e0.elapsed_time(e1)
    ms, cnt = plan.phase_ms()
    if split:
        plan.split_store(pos, vel)
    nerr = pmhip.range_errors()
    plan.close()
    phases = {n: (m / c if c else None) for n, m, c in zip(pmhip.PHASE_NAMES, ms, cnt)}
    phases["fft_shape"] = shape
    return ev_ms / args.steps, wall / args.steps * 1e3, phases, nerr, (pos, vel)


def run_adjoint_leg(pos0, vel0, args, ng, npart, variant):
    """pm_run_kdk over args.steps steps, then the full reverse sweep
    pm_run_adjoint (reverse-time recomputation, no stored trajectory) with a
    constant position cotangent; returns the time ratio and the reversal
    drift of the recovered initial positions / velocities."""
    import torch
    import pmhip

    pos = [p.clone() for p in pos0]
    vel = [v.clone() for v in vel0]
    plan = pmhip.Plan(ng, npart, pmhip.PM_FP32, pmhip.PM_GREEN_CONT_K2, 1.0)
    n = args.steps
    plan.run_kdk(pos, vel, 1, args.dt, variant)  # warm-up (allocates the binning)
    pos = [p.clone() for p in pos0]
    vel = [v.clone() for v in vel0]
    xb = [torch.full_like(p, 1e-3) for p in pos]
    vb = [torch.zeros_like(p) for p in pos]
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    torch.cuda.synchronize()
    ev[0].record(stream)
    plan.run_kdk(pos, vel, n, args.dt, variant)
    ev[1].record(stream)
    plan.run_adjoint(pos, vel, xb, vb, n, args.dt, variant)
    ev[2].record(stream)
    torch.cuda.synchronize()
    t_f, t_a = ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2])
    dx = []
    for a, b in zip(pos, pos0):
        d = a - b
        d = d - ng * torch.round(d / ng)
        dx.append(d)
    d2 = sum((d * d) for d in dx)
    vmax = max(float(v.abs().max()) for v in vel0)
    dv = max(float((a - b).abs().max()) for a, b in zip(vel, vel0))
    plan.close()
    return {"steps": n, "force_evaluations": n + 1, "forward_ms": round(t_f, 3), "adjoint_ms": round(t_a, 3),
            "ratio": round(t_a / t_f, 3), "reversal_drift_rms_cells": float(torch.sqrt(d2.mean())),
            "reversal_drift_max_cells": float(torch.sqrt(d2.max())), "reversal_dv_max_rel": dv / vmax,
            "mode": "reverse-time leapfrog recomputation, FP32 mesh, FP64 positions"}


def gpu_fields(pos0, vel0, args, ng, npart, variant, nsteps):
    """Both GPU mesh paths on the same ICs: density contrast, potential and
    per-particle force of the initial state, and the positions after nsteps
    KDK steps (pm_run_kdk); and the positions after nsteps steps of the
    split-precision state (pm_split_run_kdk, key "split")."""
    import torch
    import pmhip

    fields = {}
    plan = pmhip.Plan(ng, npart, pmhip.PM_FP32, pmhip.PM_GREEN_CONT_K2, 1.0)
    plan.split_load(pos0, vel0)
    plan.split_run_kdk(nsteps, args.dt)
    pos = [torch.empty_like(p) for p in pos0]
    vel = [torch.empty_like(v) for v in vel0]
    plan.split_store(pos, vel)
    torch.cuda.synchronize()
    plan.close()
    del vel
    fields["split"] = pos
    # the same with an FP64 mesh (u64 positions, FP64 velocities)
    plan = pmhip.Plan(ng, npart, pmhip.PM_FP64, pmhip.PM_GREEN_CONT_K2, 1.0)
    plan.split_load(pos0, vel0)
    plan.split_run_kdk(nsteps, args.dt)
    pos = [torch.empty_like(p) for p in pos0]
    vel = [torch.empty_like(v) for v in vel0]
    plan.split_store(pos, vel)
    torch.cuda.synchronize()
    plan.close()
    del vel
    fields["split64"] = pos
    for prec in (pmhip.PM_FP64, pmhip.PM_FP32):
        plan = pmhip.Plan(ng, npart, prec, pmhip.PM_GREEN_CONT_K2, 1.0)
        plan.deposit(pos0, variant)
        delta = plan.density().double() - 1.0
        plan.solve()
        phi = plan.potential().double().clone()
        g = [torch.empty_like(p) for p in pos0]
        plan.interp_force(pos0, g)
        pos = [p.clone() for p in pos0]
        vel = [v.clone() for v in vel0]
        plan.run_kdk(pos, vel, nsteps, args.dt, variant)
        torch.cuda.synchronize()
        plan.close()
        del vel
        fields[prec] = (delta, phi, g, pos)
    return fields


def host_threads():
    """The host cores this process may use, as `nproc` counts them (it
    honours OMP_NUM_THREADS and the affinity mask), and the machine's total."""
    import subprocess

    def count(cmd, fallback):
        try:
            return int(subprocess.run(cmd, capture_output=True, text=True, check=True).stdout.strip())
        except (OSError, ValueError, subprocess.CalledProcessError):
            return fallback

    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return count(["nproc"], avail), count(["nproc", "--all"], os.cpu_count() or 1)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# The oracle's one-box step against the reference binary on the same 8 cores
# of the build container (Intel Xeon @ 2.70 GHz, init_particles input,
# OMP_NUM_THREADS=8): oracle/calibrate.py gives 1.451 s/step at 256^3 and
# 12.673 s/step at 512^3; the reference (ffte1, P=1 x 8 threads, 20 steps
# after 10 warm-up) took 1.229 and 10.94 s/step on that container.
CALIBRATION = {"host": "build container, 8 cores, Intel Xeon @ 2.70 GHz", "threads": 8,
               "oracle_s_per_step": {"256": 1.451, "512": 12.673},
               "reference_s_per_step": {"256": 1.229, "512": 10.94},
               "oracle_over_reference": {"256": round(1.451 / 1.229, 3), "512": round(12.673 / 10.94, 3)},
               "how": "python oracle/calibrate.py --ng N --threads 8 --steps 3; reference: ffte1 build, "
                      "mpirun -n 1, OMP_NUM_THREADS=8, main.cpp's 20 timed steps"}


def oracle_leg(pos_dev, vel_dev, ng, args, nsteps):
    """The FP64 oracle (C/OpenMP restatement of the reference step) on the same
    ICs with OMP_NUM_THREADS = nproc.  Untimed warm-up: the initial state's
    density, potential and force (it allocates the private grids).  Timed:
    nsteps KDK steps (nsteps + 1 force evaluations, as pm_run_kdk) -- the CPU
    baseline -- whose final positions are the accuracy reference."""
    import numpy as np

    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle

    nproc, nproc_all = host_threads()
    threads = args.cpu_threads or nproc
    pyoracle.set_threads(threads)
    pos = [p.cpu().numpy().copy() for p in pos_dev]
    vel = [v.cpu().numpy().copy() for v in vel_dev]
    npart = pos[0].size
    t0 = time.perf_counter()
    delta = pyoracle.deposit_periodic(pos, ng)
    phi = pyoracle.solve(delta, pyoracle.GREEN_CONT_K2, 1.0)
    g = pyoracle.interp_force(pos, phi)
    t_warm = time.perf_counter() - t0
    t0 = time.perf_counter()
    pyoracle.run_kdk(pos, vel, ng, nsteps, args.dt, pyoracle.GREEN_CONT_K2, 1.0)
    t = time.perf_counter() - t0
    per_eval = t / (nsteps + 1)
    cpu = {"value": npart / per_eval / 1e6, "unit": "Mparticle-steps/s", "cores": threads, "kind": "port",
           "sample": f"{nsteps} KDK steps ({nsteps + 1} force evaluations, each deposit + FFT solve + "
                     f"interpolation + kick + drift) of the FP64 oracle from the same {round(npart ** (1 / 3))}^3 "
                     f"ICs after one untimed force evaluation: {t:.2f} s, {per_eval:.3f} s per evaluation, "
                     f"{threads} OpenMP threads",
           "nproc": nproc, "nproc_all": nproc_all, "cpu_model": cpu_model(), "warmup_s": round(t_warm, 3),
           "s_per_step": round(per_eval, 4), "calibration": CALIBRATION}
    return cpu, (delta.reshape(ng, ng, ng), phi, g, pos)


def accuracy_report(fields, ref, ng, nsteps, dt, dev):
    """Errors of both GPU mesh paths against the FP64 oracle (relative rms of
    the density contrast, potential and force of the initial state; final
    positions in cells after nsteps steps), of the split state's final
    positions, and of the FP32 path against the GPU FP64 path."""
    import torch
    import pmhip

    def rel(a, b):
        return float(torch.sqrt(((a - b) ** 2).mean() / (b ** 2).mean()))

    def pos_errors(x, rx):
        d2 = None
        for a, b in zip(x, rx):
            e = a - b
            e = e - ng * torch.round(e / ng)
            d2 = e * e if d2 is None else d2 + e * e
        return {"final_position_rms_cells": float(torch.sqrt(d2.mean())),
                "final_position_max_cells": float(torch.sqrt(d2.max()))}

    def errors(f, r):
        d, p, g, x = f
        rd, rp, rg, rx = r
        gnum = sum(float(((a - b) ** 2).mean()) for a, b in zip(g, rg))
        gden = sum(float((b ** 2).mean()) for b in rg)
        return {"density_rel_rms": rel(d, rd), "potential_rel_rms": rel(p, rp), "force_rel_rms": (gnum / gden) ** 0.5,
                **pos_errors(x, rx)}

    rd, rp, rg, rx = ref
    oref = (torch.from_numpy(rd).to(dev), torch.from_numpy(rp).to(dev), [torch.from_numpy(v).to(dev) for v in rg],
            [torch.from_numpy(v).to(dev) for v in rx])
    out = {"against": "FP64 oracle (C/OpenMP restatement of the reference step) on the same ICs", "steps": nsteps,
           "dt": dt, "target_rms_cells": 1e-4,
           "fp32_split": pos_errors(fields["split"], oref[3]),
           "fp64_split": pos_errors(fields["split64"], oref[3]),
           "fp32": errors(fields[pmhip.PM_FP32], oref), "fp64": errors(fields[pmhip.PM_FP64], oref),
           "fp32_vs_gpu_fp64": errors(fields[pmhip.PM_FP32], fields[pmhip.PM_FP64])}
    del oref
    return out


def run_large_adjoint_leg(args, dev):
    """512^3 particles on a 1024^3 mesh, 1.0-cell rms Zel'dovich ICs, 30 KDK
    steps with the FP32 mesh, then the full reverse sweep (no stored
    trajectory): time ratio, reversal drift of the recovered initial state
    against the FP64 initial positions, and device memory in use."""
    import torch
    import pmhip

    nrow, ng, n = args.large_nrow, args.large_ng, args.large_steps
    npart = nrow ** 3
    torch.cuda.synchronize()
    free0, total = torch.cuda.mem_get_info()
    pos0, vel0 = make_ics(nrow, ng, 1.0, args.seed + 1, dev)
    pos = [p.clone() for p in pos0]
    vel = [v.clone() for v in vel0]
    xb = [torch.full_like(p, 1e-3) for p in pos]
    vb = [torch.zeros_like(p) for p in pos]
    plan = pmhip.Plan(ng, npart, pmhip.PM_FP32, pmhip.PM_GREEN_CONT_K2, 1.0)
    variant = pmhip.PM_DEPOSIT_SORTED
    w = [p.clone() for p in pos0], [v.clone() for v in vel0]
    plan.run_kdk(w[0], w[1], 1, args.dt, variant)  # warm-up (allocates the binning)
    del w
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    torch.cuda.synchronize()
    ev[0].record(stream)
    plan.run_kdk(pos, vel, n, args.dt, variant)
    ev[1].record(stream)
    plan.run_adjoint(pos, vel, xb, vb, n, args.dt, variant)
    ev[2].record(stream)
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    t_f, t_a = ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2])
    d2 = None
    for a, b in zip(pos, pos0):
        d = a - b
        d = d - ng * torch.round(d / ng)
        d2 = d * d if d2 is None else d2 + d * d
    out = {"np": npart, "ng": ng, "steps": n, "rms_cells": 1.0, "forward_ms": round(t_f, 3),
           "adjoint_ms": round(t_a, 3), "ratio": round(t_a / t_f, 3),
           "reversal_drift_rms_cells": float(torch.sqrt(d2.mean())),
           "reversal_drift_max_cells": float(torch.sqrt(d2.max())),
           "device_memory_in_use_GB": round((total - free1) / 1e9, 2),
           "device_memory_added_GB": round((free0 - free1) / 1e9, 2),
           "mode": "FP32 mesh, FP64 positions, reverse-time leapfrog recomputation"}
    plan.close()
    del pos, vel, xb, vb, pos0, vel0, d2
    torch.cuda.empty_cache()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU's box only (got --gpus {args.gpus}); the multi-GPU slab "
              "decomposition is not built. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import torch
    import pmhip

    dev = torch.device("cuda:0")
    ng, nrow = args.ng, args.nrow
    npart = nrow ** 3
    log(f"ICs: {nrow}^3 particles, {ng}^3 mesh, rms {args.rms} cells, seed {args.seed}")
    pos0, vel0 = make_ics(nrow, ng, args.rms, args.seed, dev)

    variant = pmhip.PM_DEPOSIT_SORTED if args.deposit == "sorted" else pmhip.PM_DEPOSIT_ATOMIC
    alt_variant = pmhip.PM_DEPOSIT_ATOMIC if args.deposit == "sorted" else pmhip.PM_DEPOSIT_SORTED
    alt_name = "atomic" if args.deposit == "sorted" else "sorted"
    split = variant == pmhip.PM_DEPOSIT_SORTED
    if split:
        # the headline: split-precision positions, as the 512^3 benchmark step
        # names them (the sorted deposit; the atomic variant keeps FP64 positions)
        log("FP32-mesh leg (split-precision state, sorted deposit)")
        ms32, wall32, ph32, err32, _ = run_gpu(pmhip.PM_FP32, pos0, vel0, args, ng, npart, variant, split=True)
        log(f"FP32 split: {ms32:.3f} ms/step, phases {ph32}")
    msd = None
    if not (split and args.no_fp64_positions):
        log(f"FP32-mesh leg, FP64 positions ({args.deposit} deposit)")
        msd, walld, phd, errd, _ = run_gpu(pmhip.PM_FP32, pos0, vel0, args, ng, npart, variant)
        log(f"FP32 FP64-positions: {msd:.3f} ms/step, phases {phd}")
    if not split:
        ms32, wall32, ph32, err32 = msd, walld, dict(phd), errd
    alt = None
    if not args.no_alt_deposit:
        log(f"FP32-mesh leg ({alt_name} deposit)")
        ms_a, _, ph_a, _, _ = run_gpu(pmhip.PM_FP32, pos0, vel0, args, ng, npart, alt_variant)
        alt = {"deposit": alt_name, "value": round(npart / (ms_a * 1e-3) / 1e6, 2), "ms_per_step": round(ms_a, 4),
               "deposit_ms": round(ph_a["deposit"], 4),
               "deposit_hbm": deposit_traffic(alt_name, ng, npart, ph_a["deposit"])}
        log(f"FP32 {alt_name}: {ms_a:.3f} ms/step")
    res64 = None
    if not args.no_fp64:
        log("FP64-mesh leg")
        ms64, wall64, ph64, err64, _ = run_gpu(pmhip.PM_FP64, pos0, vel0, args, ng, npart, variant)
        log(f"FP64: {ms64:.3f} ms/step, phases {ph64}")
        res64 = {"value": npart / (ms64 * 1e-3) / 1e6, "ms_per_step": ms64, "phases_ms": ph64}
        log("FP64-mesh leg, split state (u64 positions, FP64 velocities)")
        ms64s, _, ph64s, err64s, _ = run_gpu(pmhip.PM_FP64, pos0, vel0, args, ng, npart, variant, split=True)
        log(f"FP64 split: {ms64s:.3f} ms/step, phases {ph64s}")
        res64["split"] = {"value": npart / (ms64s * 1e-3) / 1e6, "ms_per_step": ms64s, "phases_ms": ph64s,
                          "range_errors": err64s}

    fields = None
    if not args.no_accuracy:
        log(f"accuracy leg: both GPU mesh paths, initial fields and {args.accuracy_steps} KDK steps")
        fields = gpu_fields(pos0, vel0, args, ng, npart, variant, args.accuracy_steps)

    adj = None
    if not args.no_adjoint:
        log("adjoint leg")
        adj = run_adjoint_leg(pos0, vel0, args, ng, npart, variant)
        log(f"adjoint: {adj}")

    large = None
    if not args.no_large_adjoint:
        log(f"large adjoint leg ({args.large_nrow}^3 particles, {args.large_ng}^3 mesh, {args.large_steps} steps)")
        large = run_large_adjoint_leg(args, dev)
        log(f"large adjoint: {large}")

    if args.copy_calibration:
        # a plain streaming copy of 1 GiB (read 1 GiB, write 1 GiB) for PMC byte-counter scaling
        a = torch.empty(1 << 28, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        a.fill_(1.0)
        for _ in range(3):
            torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()

    # the device's measured streaming ceiling (1 GiB read + 1 GiB written, one
    # 16-byte load and store a thread), beside the 8 TB/s nameplate
    copy_gbs = pmhip.copy_rate(1 << 30, 20)
    log(f"device copy rate: {copy_gbs:.0f} GB/s (read + write)")

    shape32 = ph32.pop("fft_shape")
    if msd is not None:
        phd.pop("fft_shape", None)
    shape64 = res64["phases_ms"].pop("fft_shape") if res64 else None
    if res64:
        res64["split"]["phases_ms"].pop("fft_shape", None)
    sorted_dep = variant == pmhip.PM_DEPOSIT_SORTED

    def fp32_roofline(ph, is_split):
        """Roofline of the dominant kernel of an FP32-mesh step: its
        algorithmic bytes (with the staging record; beside them without
        it) over its average launch time, and its PMC traffic."""
        nb = algorithmic_bytes(ng, npart, pmhip.PM_FP32, sorted_dep, is_split)
        dom = max((k for k in ph if ph[k]), key=lambda k: ph[k])
        achieved = nb[dom] / (ph[dom] * 1e-3) / 1e9
        kname = (SPLIT_PHASE_KERNELS if is_split else PHASE_KERNELS).get(dom) if sorted_dep else None
        traffic = measured_traffic(kname, ng, npart) if kname else None
        roof = {"bound": "hbm", "kernel": kname or dom, "phase": dom, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "algorithmic_bytes": nb[dom], "avg_ms": round(ph[dom], 4),
                "traffic_source": f"profiles/pmc_traffic_{ng}.txt" if traffic else None}
        # the same kernel priced without its staging record (the step's
        # compulsory bytes only: b N_c + 2 x state N_p for the interpolation,
        # 24 N_p + b N_c for the deposit)
        unstaged = algorithmic_bytes(ng, npart, pmhip.PM_FP32, False, is_split)[dom]
        roof["algorithmic_bytes_unstaged"] = unstaged
        roof["frac_unstaged"] = round(unstaged / (ph[dom] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        roof["copy_ceiling"] = round(copy_gbs, 1)
        roof["frac_of_copy"] = round(achieved / copy_gbs, 4)
        return roof, sum(nb.values())

    roofline, step_bytes = fp32_roofline(ph32, split)
    roofline_d, step_bytes_d = fp32_roofline(phd, False) if msd is not None else (None, None)
    roofline64 = None
    if res64:
        ph64 = res64["phases_ms"]
        nb64 = algorithmic_bytes(ng, npart, pmhip.PM_FP64, variant == pmhip.PM_DEPOSIT_SORTED)
        dom64 = max((k for k in ph64 if ph64[k]), key=lambda k: ph64[k])
        ach64 = nb64[dom64] / (ph64[dom64] * 1e-3) / 1e9
        k64 = FP64_INTERP_KERNEL if (dom64 == "interp_kick_drift" and variant == pmhip.PM_DEPOSIT_SORTED) else None
        tr64 = measured_traffic_fp64(k64, ng) if k64 else None
        roofline64 = {"bound": "hbm", "kernel": k64 or dom64, "phase": dom64, "achieved": round(ach64, 1),
                      "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach64 / HBM_PEAK_GBS, 4),
                      "traffic": tr64, "algorithmic_bytes": nb64[dom64], "avg_ms": round(ph64[dom64], 4),
                      "traffic_source": f"profiles/pmc_traffic_{ng}_fp64_adjoint.txt" if tr64 else None}
        un64 = algorithmic_bytes(ng, npart, pmhip.PM_FP64, False)[dom64]
        roofline64["algorithmic_bytes_unstaged"] = un64
        roofline64["frac_unstaged"] = round(un64 / (ph64[dom64] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        roofline64["copy_ceiling"] = round(copy_gbs, 1)
        roofline64["frac_of_copy"] = round(ach64 / copy_gbs, 4)
        # the FP64 mesh with the split state: its interpolation streams 96 B
        # a particle (u64 position and FP64 velocity triples read and
        # written) and reads the 8 B potential
        ph64s = res64["split"]["phases_ms"]
        nb64s = algorithmic_bytes(ng, npart, pmhip.PM_FP64, sorted_dep, True)
        dom64s = max((k for k in ph64s if ph64s[k]), key=lambda k: ph64s[k])
        ach64s = nb64s[dom64s] / (ph64s[dom64s] * 1e-3) / 1e9
        k64s = FP64_SPLIT_INTERP_KERNEL if dom64s == "interp_kick_drift" else None
        tr64s = measured_traffic_fp64(k64s, ng) if k64s else None
        res64["split"]["roofline"] = {
            "bound": "hbm", "kernel": k64s or dom64s, "phase": dom64s, "achieved": round(ach64s, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach64s / HBM_PEAK_GBS, 4), "traffic": tr64s,
            "algorithmic_bytes": nb64s[dom64s], "avg_ms": round(ph64s[dom64s], 4),
            "traffic_source": f"profiles/pmc_traffic_{ng}_fp64_adjoint.txt" if tr64s else None,
            "copy_ceiling": round(copy_gbs, 1), "frac_of_copy": round(ach64s / copy_gbs, 4)}
        res64["split"]["step_algorithmic_GB"] = round(sum(nb64s.values()) / 1e9, 3)
    cpu, acc = None, None
    if not args.no_cpu:
        log(f"CPU baseline leg: the FP64 oracle, {args.accuracy_steps} KDK steps on the host cores")
        cpu, ref = oracle_leg(pos0, vel0, ng, args, args.accuracy_steps)
        log(f"CPU: {cpu['value']:.2f} Mparticle-steps/s ({cpu['cores']} threads)")
        if fields is not None:
            acc = accuracy_report(fields, ref, ng, args.accuracy_steps, args.dt, dev)
            log(f"accuracy: {acc}")
        del ref
    fields = None
    value = npart / (ms32 * 1e-3) / 1e6
    out = {
        "metric": "Mparticle-steps/s at 512^3 (FP32 vs FP64 PM, 1 GPU)",
        "value": round(value, 2),
        "unit": "Mparticle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms32, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f32 mesh/FFT/force/velocities, split positions (u32 cell + fraction, 2^-23 cell)" if split
                  else "f32 mesh/FFT/force, f64 positions/velocities"),
        "data": "synthetic Zel'dovich lattice ICs generated on the GPU",
        "config": {"workload": f"PM KDK step, {nrow}^3 particles on a {ng}^3 periodic mesh, Zel'dovich k^-2 ICs "
                               f"{args.rms}-cell rms, dt={args.dt}, {args.deposit} CIC deposit, 5-pass FFT, "
                               f"CONT_K2 Green" + (", split-precision positions" if split else ""),
                   "ng": ng, "np": npart, "dt": args.dt, "rms_cells": args.rms},
        "phases_ms": {k: (round(v, 4) if v else v) for k, v in ph32.items()},
        "deposit_hbm": deposit_traffic("split" if split else args.deposit, ng, npart, ph32.get("deposit")),
        "step_algorithmic_GB": round(step_bytes / 1e9, 3),
        "step_achieved_GBs": round(step_bytes / (ms32 * 1e-3) / 1e9, 1),
        "wall_ms_per_step": round(wall32, 4),
        "fp32_fp64_positions": ({"value": round(npart / (msd * 1e-3) / 1e6, 2), "ms_per_step": round(msd, 4),
                                 "phases_ms": {k: (round(v, 4) if v else v) for k, v in phd.items()},
                                 "step_algorithmic_GB": round(step_bytes_d / 1e9, 3),
                                 "roofline": roofline_d, "range_errors": errd} if split and msd is not None
                                else None),
        "fp64": ({"value": round(res64["value"], 2), "ms_per_step": round(res64["ms_per_step"], 4),
                  "phases_ms": {k: (round(v, 4) if v else v) for k, v in res64["phases_ms"].items()},
                  "roofline": roofline64}
                 if res64 else None),
        "fp64_split": ({"value": round(res64["split"]["value"], 2),
                        "ms_per_step": round(res64["split"]["ms_per_step"], 4),
                        "phases_ms": {k: (round(v, 4) if v else v) for k, v in res64["split"]["phases_ms"].items()},
                        "step_algorithmic_GB": res64["split"]["step_algorithmic_GB"],
                        "roofline": res64["split"]["roofline"], "range_errors": res64["split"]["range_errors"]}
                       if res64 else None),
        "fft_shape": {"fp32": shape32, "fp64": shape64, "tuned": not args.no_tune,
                      "slots": ["z_fwd", "y_fwd", "x_green", "y_inv", "z_inv"]},
        "alt_deposit": alt,
        "accuracy": acc,
        "adjoint_forward_ratio": adj["ratio"] if adj else None,
        "adjoint": adj,
        "large_adjoint": large,
        "range_errors": err32,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["vs_cpu_baseline"] = round(value / cpu["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
