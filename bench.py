#!/usr/bin/env python3
"""Throughput of the explicit 3D-3V PIC-MCC step on one MI355X.

Workload: 128^3-cell grounded box, 64 particles/cell/species of e- and Ar+
(268,435,456 particles), uniform Bz = 0.01 T, argon background at 1 Pa/300 K
with the analytic cross sections, null-collision MCC every 10 steps (Philox,
parallel), cell re-sort every 20 steps, FP64 particles, FP64 multigrid
Poisson solve to 1e-8.  Particles are synthetic (device loader).

Prints one JSON line: particle-steps/s of the whole step, the roofline of the
fused particle kernel (HIP events around its launches, on the stream it runs
on) and a CPU baseline from the oracle timed on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_PARTICLE_STEP = 104   # FP64 fused gather+push+BC+deposit, state read and written once
COPY_PROBE_BYTES = 1 << 31


def measured_traffic(kernel, args):
    """HBM bytes per launch of the fused particle kernel on this workload, from
    the committed rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE summary (counter
    readings scaled by what they read for a plain copy of known size); None
    when no summary covers this kernel and workload."""
    if args.n != 128 or args.ppc != 64:
        return None
    path = os.path.join(ROOT, "profiles", f"{kernel}_128cube_64ppc_traffic.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def _cpu_chunk(job):
    """One host core's share of the particle sample: the oracle's Boris push,
    wall BC and deposit (into its own density array, as the reference's
    per-thread np_thread) on `count` particles; returns the seconds spent."""
    import numpy as np
    from oracle import oracle as O
    (n, h, bcnd, kq, Bz, q, mass, dt, Nm, ngrid, count, seed, barrier) = job
    rng = np.random.default_rng(seed)
    L = np.array([n[d] * h[d] for d in range(3)])
    pos = rng.random((count, 3)) * L
    vel = 1e6 * rng.standard_normal((count, 3))
    p = O.Particles(count + 16).set(pos[:, 0], pos[:, 1], pos[:, 2], vel[:, 0], vel[:, 1],
                                    vel[:, 2])
    E4 = np.zeros(bcnd.shape + (4,))
    E4[..., :3] = 1e3 * rng.standard_normal(bcnd.shape + (3,))
    barrier.wait()
    t0 = time.perf_counter()
    O.move_boris(n, h, p, E4, 1, (0.0, 0.0, Bz), None, q, mass, dt)
    O.apply_bc(n, h, p, bcnd, 0, 0, 1, -1, Nm * q, mass, Nm, ngrid)
    O.deposit(n, h, p, kq, Nm)
    return time.perf_counter() - t0


def _cpu_init(barrier):
    global _BARRIER
    _BARRIER = barrier


def _cpu_chunk_in_pool(job):
    return _cpu_chunk(job + (_BARRIER,))


def host_cores():
    """Cores this process may run on: its CPU affinity, capped by the cgroup
    CPU quota when one is set (a container's share of a larger host)."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            quota, period = f.read().split()[:2]
        if quota != "max":
            cores = min(cores, max(1, int(int(quota) // int(period))))
    except (OSError, ValueError):
        pass
    return max(1, cores)


def cpu_baseline(case, sample_particles=32_000_000):
    """The oracle (C restatement of the reference) on every host core this
    process may use, one process per core (spawned, so none of them touches
    the GPU): each pushes, wall-tests and deposits its share of a bounded
    particle sample on the 128^3 grid at the same moment; one warm MG solve
    with the multigrid loops on OpenMP threads over the same cores; E + rho
    at 128^3 on one core; composed into particle-steps/s of the full
    workload."""
    import multiprocessing as mp
    import numpy as np
    from oracle import oracle as O
    O.build()
    n, h = case.n, case.h
    cores = host_cores()
    sp = case.species[0]
    bc, kq = case.bcnd, case.kq()
    m = sample_particles
    per = [m // cores + (1 if k < m % cores else 0) for k in range(cores)]
    ctx = mp.get_context("spawn")
    barrier = ctx.Barrier(cores)
    jobs = [(n, h, bc, kq, case.B_uniform[2], sp.charge, sp.mass, case.dt, case.Nm[0],
             case.ngrid, c, 7 + k) for k, c in enumerate(per)]
    with ctx.Pool(cores, initializer=_cpu_init, initargs=(barrier,)) as pool:
        secs = pool.map(_cpu_chunk_in_pool, jobs)
    t_part = max(secs) / m              # seconds per particle on `cores` cores together
    # grid phase: rho from a smooth density, then one solve warm-started from its own solution
    os.environ["OMP_NUM_THREADS"] = str(cores)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    rng = np.random.default_rng(7)
    rho = np.zeros(bc.shape)
    rho[bc == -1] = 1e-5 * rng.standard_normal(int((bc == -1).sum()))
    phi = np.zeros(bc.shape)
    O.pdesolver_threaded(n, h, phi, rho, bc, ng=case.ng, eps=case.eps, omega=case.omega)
    rho *= 1.001
    t0 = time.perf_counter()
    O.pdesolver_threaded(n, h, phi, rho, bc, ng=case.ng, eps=case.eps, omega=case.omega)
    t_mg = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.efield(n, h, phi, bc)
    npt = np.zeros((2,) + bc.shape)
    O.build_rho(n, O.reduce_density(n, bc, npt), [-1.6e-19, 1.6e-19])
    t_grid = time.perf_counter() - t0
    N = 2 * case.np_cell * n[0] * n[1] * n[2]
    step = N * t_part + t_mg + t_grid
    return {"value": N / step, "unit": "particle-steps/s", "cores": cores, "kind": "port",
            "sample": (f"oracle on {cores} host cores (one process each, concurrent): Boris "
                       f"push+wall BC+deposit of {m} particles on the {n[0]}^3 grid "
                       f"({t_part * 1e9:.1f} ns/particle for the cores together), one warm MG "
                       f"solve on {cores} OpenMP threads ({t_mg * 1e3:.0f} ms) and E+rho "
                       f"({t_grid * 1e3:.0f} ms) on one core, composed to the {N}-particle step")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--ppc", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--resort-electrons", type=int, default=4,
                    help="production path: also re-sort electrons by cell every k steps "
                         "(0: only at the reference's sort cadence)")
    ap.add_argument("--exact-order", action="store_true",
                    help="compact every step by look-back (the reference's particle order "
                         "at every step) instead of in place with holes removed at each sort")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the timed steps, copy 2 GiB device to device once: the "
                         "calibration kernel for rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    import torch
    from essaitemporaire_lepic_module_amd import _native
    from essaitemporaire_lepic_module_amd import config as cfg
    from essaitemporaire_lepic_module_amd.simulation import Simulation

    case = cfg.cfg2_pic_mcc(n=args.n, np_cell=args.ppc)
    torch.cuda.init()
    resort = None if args.exact_order or args.resort_electrons <= 0 else \
        [args.resort_electrons] + [0] * (len(case.tracked) - 1)
    sim = Simulation(case, "cuda:0", fused=True, loader="device", mcc_mode=_native.MCC_PHILOX,
                     capacity_factor=1.25, lazy=not args.exact_order, resort_every=resort)
    istep = 1
    for _ in range(args.warmup):
        sim.advance_one_step(istep)
        istep += 1
    n_before = sim.particle_count()
    torch.cuda.synchronize()
    # timed region: events only around the fused particle kernel's launches;
    # live particles = slots (host) - holes (device, summed without a sync)
    sim.timers = {}
    sim.timed_phases = {"push_bc_deposit"}
    holes = torch.zeros_like(sim.lost_acc)
    slots = 0
    mg_cycles = []
    marks = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        marks[k].record()
        slots += sum(p.n for p in sim.parts)
        holes += sim.lost_acc
        sim.advance_one_step(istep)
        mg_cycles.append(sim.last_mg["cycles"])
        istep += 1
    marks[-1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = sorted(a.elapsed_time(b) for a, b in zip(marks[:-1], marks[1:]))
    particle_steps = slots - int(holes.sum().item())
    kernel_phases = sim.phase_times()
    # diagnostic pass (not timed for the value): every phase bracketed by events
    sim.timers = {}
    sim.timed_phases = None
    for _ in range(args.steps):
        sim.advance_one_step(istep)
        istep += 1
    torch.cuda.synchronize()
    if args.copy_probe:
        a = torch.ones(COPY_PROBE_BYTES // 8, dtype=torch.float64, device="cuda:0")
        torch.empty_like(a).copy_(a)
        torch.cuda.synchronize()
        del a
    phases = sim.phase_times()
    fused_ms, fused_launches, fused_units = kernel_phases.get("push_bc_deposit", (0.0, 0, 0))
    kernel = "k_push_bc_deposit" if args.exact_order else "k_push"
    achieved = BYTES_PER_PARTICLE_STEP * fused_units / (fused_ms * 1e-3) / 1e9 if fused_ms else 0.0
    mg_ms, mg_calls, _ = phases.get("poisson", (0.0, 1, 0))
    line = {
        "metric": "particle-steps/s of the 128^3 x 64 ppc x 2 species PIC-MCC step",
        "value": particle_steps / wall,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "step_ms_median": step_ms[len(step_ms) // 2],
        "step_ms_p90": step_ms[min(len(step_ms) - 1, (9 * len(step_ms)) // 10)],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (device Philox loader, Maxwellian e- 3 eV / Ar+ 0.03 eV)",
        "config": {"workload": "PIC-MCC at scale: 128^3 cells, 64 ppc x 2 species, Bz=0.01 T, "
                               "Ar 1 Pa, MCC every 10, sort every 20, FP64",
                   "cells": args.n ** 3, "particles": n_before, "ppc": args.ppc,
                   "electron_resort_every": 0 if resort is None else resort[0],
                   "parallelism": "1 GPU"},
        "mg_solve_ms": mg_ms / max(mg_calls, 1),
        "mg_vcycles_per_solve": sum(mg_cycles) / max(len(mg_cycles), 1),
        "phase_ms_per_step": {k: v[0] / args.steps for k, v in phases.items()},
        "phase_timing": "second pass of the same number of steps, every phase between HIP "
                        "events (the timed pass brackets only the fused particle kernel)",
        "roofline": {"bound": "hbm", "kernel": kernel,
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(kernel, args),
                     "bytes_per_particle": BYTES_PER_PARTICLE_STEP,
                     "algorithmic_bytes_per_launch":
                         BYTES_PER_PARTICLE_STEP * fused_units / max(fused_launches, 1),
                     "avg_launch_ms": fused_ms / max(fused_launches, 1)},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(case)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
