#!/usr/bin/env python3
"""Throughput of the explicit 3D-3V PIC-MCC step on one MI355X.

Workload: 128^3-cell grounded box, 64 particles/cell/species of e- and Ar+
(268,435,456 particles), uniform Bz = 0.01 T, argon background at 1 Pa/300 K
with the analytic cross sections, null-collision MCC every 10 steps (Philox,
parallel), cell re-sort every 20 steps, FP64 multigrid Poisson solve to 1e-8.
Particles are synthetic (device loader).

The headline is the step in its precision form, as the workload specifies
it: FP32 particles (int32 cell + FP32 in-cell offsets + FP32 velocities),
FP32 gather and push, int64 fixed-point deposit, FP64 fields.  The FP64 step
is measured beside it in the same run; so are the mixed-precision Poisson
solves (FP32 red-black V-cycles inside FP64 refinement, and FP64 conjugate
gradients preconditioned with them) against the FP64 multigrid on their own
256^3 problem, the combined annular-channel run at 256^3 (392M and 1.07B
particles) and the forward + adjoint costs of the differentiable runs.

Prints one JSON line: particle-steps/s of the whole step, the roofline of the
fused particle kernel (HIP events around its launches, on the stream it runs
on) and a CPU baseline from the oracle timed on this host.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_PARTICLE_STEP = 104   # FP64 fused gather+push+BC+deposit, state read and written once
BYTES_PER_PARTICLE_STEP_F32 = 56   # FP32: int32 cell + 3 FP32 offsets + 3 FP32 velocities, r + w
COPY_PROBE_BYTES = 1 << 31
# FP64 multigrid, per node and executed sweep: u read and written, b and the
# int32 label read (28 B), and the sweep's residual pass (u, b, label: 20 B);
# restriction and prolongation not counted
MG_BYTES_PER_NODE_SWEEP = 48
SOLVES = {"fp64": "the FP64 multigrid",
          "pcg64": "FP64 conjugate gradients preconditioned with one FP64 V-cycle",
          "pcg": "FP64 conjugate gradients preconditioned with one FP32 V-cycle"}


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


def _step_bench(args, precision):
    """W warm-up steps, then K steps timed by wall clock between two device
    synchronisations (HIP events only around the fused particle kernel),
    then K more steps with every phase between events (not part of value).
    precision "f32": the step's precision form (FP32 particles in cell +
    offset layout, fixed-point deposit, FP64 fields); "f64": the FP64 step."""
    import torch
    from essaitemporaire_lepic_module_amd import _native
    from essaitemporaire_lepic_module_amd import config as cfg
    from essaitemporaire_lepic_module_amd.simulation import Simulation, Simulation32

    case = cfg.cfg2_pic_mcc(n=args.n, np_cell=args.ppc)
    resort = None if args.exact_order or args.resort_electrons <= 0 else \
        [args.resort_electrons] + [0] * (len(case.tracked) - 1)
    if precision == "f32":
        sim = Simulation32(case, "cuda:0", capacity_factor=1.25, resort_every=resort,
                           poisson=args.step_poisson)
        kernel, bpp = "k_push32", BYTES_PER_PARTICLE_STEP_F32
    else:
        sim = Simulation(case, "cuda:0", fused=True, loader="device",
                         mcc_mode=_native.MCC_PHILOX, capacity_factor=1.25,
                         lazy=not args.exact_order, resort_every=resort)
        kernel = "k_push_bc_deposit" if args.exact_order else "k_push"
        bpp = BYTES_PER_PARTICLE_STEP
    istep = 1
    # the warm-up steps run the timed loop's own bookkeeping (events, the
    # device hole count), so that nothing in it is used for the first time
    # inside the timed region
    sim.timed_phases = {"push_bc_deposit"}
    marks = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    holes = torch.zeros_like(sim.lost_acc)
    for k in range(args.warmup):
        sim.timers = {}
        marks[k % args.steps].record()
        holes += sim.lost_acc
        sim.advance_one_step(istep)
        istep += 1
    for m in marks:
        m.record()
    sim.phase_times()
    n_before = sim.particle_count()
    first_timed = istep
    torch.cuda.synchronize()
    sim.timers = {}
    holes.zero_()
    slots = 0
    mg_cycles = []
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
    step_each = [a.elapsed_time(b) for a, b in zip(marks[:-1], marks[1:])]
    step_ms = sorted(step_each)
    particle_steps = slots - int(holes.sum().item())
    kernel_phases = sim.phase_times()
    sim.timers = {}
    sim.timed_phases = None
    mg_node_sweeps = 0.0
    for _ in range(args.steps):
        sim.advance_one_step(istep)
        istep += 1
        # sweeps per level of this solve (each with its residual pass) x the
        # level's nodes, (n / 2^l + 1)^3
        mg_node_sweeps += sum(sw * ((args.n >> l) + 1) ** 3
                              for l, sw in enumerate(sim.last_mg["sweeps"]))
    torch.cuda.synchronize()
    phases = sim.phase_times()
    if precision == "f32":
        sim.check_escaped()
    fused_ms, fused_launches, fused_units = kernel_phases.get("push_bc_deposit", (0.0, 0, 0))
    achieved = bpp * fused_units / (fused_ms * 1e-3) / 1e9 if fused_ms else 0.0
    mg_ms, mg_calls, _ = phases.get("poisson", (0.0, 1, 0))
    mg_bytes = MG_BYTES_PER_NODE_SWEEP * mg_node_sweeps
    mg_gbs = mg_bytes / (mg_ms * 1e-3) / 1e9 if mg_ms else 0.0
    out = {
        "value": particle_steps / wall,
        "ms_per_step": wall * 1e3 / args.steps,
        "step_ms_median": step_ms[len(step_ms) // 2],
        "step_ms_p90": step_ms[min(len(step_ms) - 1, (9 * len(step_ms)) // 10)],
        "step_ms_each": [round(v, 3) for v in step_each],
        "first_timed_step": first_timed,
        "particles": n_before,
        "electron_resort_every": 0 if resort is None else resort[0],
        "mg_solve_ms": mg_ms / max(mg_calls, 1),
        "mg_vcycles_per_solve": sum(mg_cycles) / max(len(mg_cycles), 1),
        "mg_roofline": {"bound": "hbm", "bytes_per_node_sweep": MG_BYTES_PER_NODE_SWEEP,
                        "algorithmic_GB_per_solve": mg_bytes / max(mg_calls, 1) / 1e9,
                        "achieved": mg_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": mg_gbs / HBM_PEAK_GBS},
        "phase_ms_per_step": {k: v[0] / args.steps for k, v in phases.items()},
        "roofline": {"bound": "hbm", "kernel": kernel,
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(kernel, args),
                     "bytes_per_particle": bpp,
                     "algorithmic_bytes_per_launch": bpp * fused_units / max(fused_launches, 1),
                     "avg_launch_ms": fused_ms / max(fused_launches, 1)},
    }
    del sim
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return out


def _poisson_cfg3(n):
    """The mixed-precision Poisson workload: n^3 cells, six faces biased at
    U[-100, 100] V, 64 Gaussian blobs of charge, solved to 1e-10 (L1 / alpha)
    from phi = 0 inside, once with the reference's FP64 multigrid, once
    with FP32 red-black V-cycles inside FP64 refinement and once with FP64
    conjugate gradients preconditioned with one FP32 V-cycle; each solve timed
    after one untimed solve of the same problem (graph capture).  Errors of
    phi and E (max abs over nodes, relative to the FP64 solution's max)."""
    import torch
    from essaitemporaire_lepic_module_amd import config as cfg
    from essaitemporaire_lepic_module_amd import fields as F
    n3, h, bc, phi0, rhs = cfg.cfg3_poisson(n=n)
    # preconditioned conjugate gradients with one FP32 V-cycle per iteration
    pcg = lambda *a: F.solve_poisson_pcg(*a[:9])
    dev = "cuda:0"
    bcg = torch.as_tensor(bc, device=dev)
    rg = torch.as_tensor(rhs, device=dev)
    ng = max(2, int(round(math.log2(n))))
    res = {}
    phis = {}
    for name, solve in (("fp64", F.solve_poisson_legacy), ("mixed", F.solve_poisson_mixed),
                        ("pcg", pcg)):
        for rep in range(2):
            phi = torch.as_tensor(phi0, device=dev).clone()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            info = solve(phi, bcg, rg, h, n3, 20000, 1e-10, 1.4, ng, 0, 0)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
        res[name] = {"ms_per_solve": dt * 1e3, "cycles": info["cycles"], "res": info["res"]}
        phis[name] = phi
    E = {}
    for name, phi in phis.items():
        E4 = torch.zeros(phi.shape + (4,), dtype=torch.float64, device=dev)
        F.calc_Efield_modular(n3, h, phi, E4, bcg)
        E[name] = E4
    pmax = float(phis["fp64"].abs().max())
    emax = float(E["fp64"][..., :3].abs().max())
    res["phi_err_rel"] = float((phis["mixed"] - phis["fp64"]).abs().max()) / pmax
    res["E_err_rel"] = float((E["mixed"][..., :3] - E["fp64"][..., :3]).abs().max()) / emax
    res["speedup"] = res["fp64"]["ms_per_solve"] / res["mixed"]["ms_per_solve"]
    res["pcg"]["phi_err_rel"] = float((phis["pcg"] - phis["fp64"]).abs().max()) / pmax
    res["pcg"]["E_err_rel"] = float((E["pcg"][..., :3] - E["fp64"][..., :3]).abs().max()) / emax
    res["pcg"]["speedup"] = res["fp64"]["ms_per_solve"] / res["pcg"]["ms_per_solve"]
    res["workload"] = (f"{n}^3 cells, Dirichlet faces U[-100,100] V, 64 Gaussian blobs, "
                       f"tolerance 1e-10, {ng} levels, from phi = 0")
    del phis, E, bcg, rg
    torch.cuda.empty_cache()
    return res


def _adjoint_cfg4(steps, n=64, ppc=8, checkpoint=20, poisson="fp64"):
    """The differentiable run: n^3 cells x ppc per species, one wall biased
    at 20 V, Bz = 0.005 T, argon MCC every 10 steps frozen in the backward,
    `steps` steps with checkpoints every `checkpoint`, the particles re-sorted
    by cell every 20, the forward and adjoint Poisson solves by `poisson`
    (DiffPIC's).  Times the plain forward (no checkpoints; the faster of two
    runs), the forward that keeps checkpoints and the backward (recompute +
    adjoint kernels), after an untimed forward + backward of the same run."""
    import torch
    import numpy as np
    from essaitemporaire_lepic_module_amd import config as cfg
    from essaitemporaire_lepic_module_amd.adjoint import DiffPIC
    case = cfg.cfg4_differentiable(n=n, np_cell=ppc, steps=steps)
    rng = np.random.default_rng(0)
    N = ppc * n ** 3
    L = np.array([n * case.h[d] for d in range(3)])
    pos = rng.random((N, 3)) * L
    parts = []
    for sp in case.tracked:
        vel = rng.standard_normal((N, 3)) * cfg.thermal_speed(sp.T_eV, sp.mass) / math.sqrt(2.0)
        parts.append({"x": pos[:, 0], "y": pos[:, 1], "z": pos[:, 2],
                      "vx": vel[:, 0], "vy": vel[:, 1], "vz": vel[:, 2]})
    sim = DiffPIC(case, parts, "cuda:0", checkpoint=checkpoint, poisson=poisson)
    sim.forward(20.0, 0.005, steps, keep=True)
    sim.backward()

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, r

    # the plain forward timed twice, the faster kept (the first full-length
    # run can carry first-use costs the short warm-up did not)
    t_plain, loss = min(timed(lambda: sim.forward(20.0, 0.005, steps, keep=False))
                        for _ in range(2))
    t_keep, _ = timed(lambda: sim.forward(20.0, 0.005, steps, keep=True))
    t_back, (dV, dB) = timed(sim.backward)
    ncoll = sum(int(ev["idx"].numel()) for evs in sim.mcc_events.values() for ev in evs)
    out = {"forward_ms": t_plain * 1e3, "forward_checkpointed_ms": t_keep * 1e3,
           "backward_ms": t_back * 1e3,
           "adjoint_over_forward": (t_keep + t_back) / t_plain,
           "loss": loss, "dL_dV": dV, "dL_dB0": dB, "collided": ncoll,
           "workload": (f"{n}^3 cells, {ppc} ppc x 2 species, wall biased at 20 V, Bz 0.005 T, "
                        f"argon MCC every 10 steps (events frozen in the backward), {steps} "
                        f"steps, checkpoint every {checkpoint}, re-sort every 20, FP64, forward "
                        f"Poisson solves by {SOLVES[poisson]}, adjoint solves by "
                        f"{SOLVES['pcg']}")}
    del sim
    torch.cuda.empty_cache()
    return out


def _adjoint_cfg5(n, ppc=32, steps=500, checkpoint=20, seed=0, poisson="pcg", resort=20):
    """The combined run's forward + adjoint of the ion current into the anode
    with respect to the anode bias, on an n^3 grid of the same geometry (the
    annular channel, anode at 300 V, radial Br(z) map, argon MCC every 10
    steps frozen in the backward), FP64, `steps` steps with checkpoints
    every `checkpoint`.  Particles: ppc per species in the loadable cells,
    uniform positions and Maxwellian velocities from a seeded host draw.
    Times the plain forward (the faster of two runs), the forward that keeps
    checkpoints and its backward, after an untimed 20-step forward +
    backward."""
    import torch
    import numpy as np
    from essaitemporaire_lepic_module_amd import config as cfg
    from essaitemporaire_lepic_module_amd.adjoint import DiffPIC
    case = cfg.cfg5_combined(n=n, np_cell=ppc, steps=steps)
    rng = np.random.default_rng(seed)
    h = np.array(case.h)
    pl = np.asarray(case.bcnd) == -1
    load = np.zeros((n, n, n), bool)
    for dk in (1, 2):
        for dj in (1, 2):
            for di in (1, 2):
                load |= pl[dk:dk + n, dj:dj + n, di:di + n]
    cells = np.flatnonzero(load.reshape(-1))
    ci = np.repeat(cells, ppc)
    pos = np.stack([ci % n, (ci // n) % n, ci // (n * n)], 1).astype(np.float64)
    pos = (pos + rng.random(pos.shape)) * h
    parts = []
    for sp in case.tracked:
        vel = rng.standard_normal(pos.shape) * cfg.thermal_speed(sp.T_eV, sp.mass) / math.sqrt(2.0)
        parts.append({"x": pos[:, 0], "y": pos[:, 1], "z": pos[:, 2],
                      "vx": vel[:, 0], "vy": vel[:, 1], "vz": vel[:, 2]})
    del pos, ci
    sim = DiffPIC(case, parts, "cuda:0", checkpoint=checkpoint, poisson=poisson, resort=resort)
    del parts
    sim.forward(300.0, None, 20, keep=True)
    sim.backward()

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, r

    t_plain, loss = min(timed(lambda: sim.forward(300.0, None, steps, keep=False))
                        for _ in range(2))
    torch.cuda.reset_peak_memory_stats()
    t_keep, _ = timed(lambda: sim.forward(300.0, None, steps, keep=True))
    t_back, (dV, _) = timed(sim.backward)
    ncoll = sum(int(ev["idx"].numel()) for evs in sim.mcc_events.values() for ev in evs)
    out = {"forward_ms": t_plain * 1e3, "forward_checkpointed_ms": t_keep * 1e3,
           "backward_ms": t_back * 1e3,
           "adjoint_over_forward": (t_keep + t_back) / t_plain,
           "loss_A": loss, "dL_dV_A_per_V": dV, "collided": ncoll,
           "particles": 2 * int(len(cells)) * ppc,
           "peak_device_GB": torch.cuda.max_memory_allocated() / 1e9,
           "workload": (f"{n}^3 cells of the combined run's geometry (annular channel, anode "
                        f"300 V, radial Br(z) map), {ppc} ppc x 2 species in the loadable "
                        f"cells, argon MCC every 10 steps (frozen in the backward), {steps} "
                        f"steps, checkpoint every {checkpoint}, re-sort every 20, FP64, forward "
                        f"Poisson solves by {SOLVES[poisson]}, adjoint solves by "
                        f"{SOLVES['pcg']}, to 1e-10; loss: the mean ion current into the anode "
                        "through z = h3")}
    del sim
    torch.cuda.empty_cache()
    return out


def _combined_cfg5(n, ppc, steps=10, warmup=3, resort=4):
    """The combined large run (cfg5 forward): n^3 cells, the annular channel
    with the anode at 300 V, the radial Br(z) map, argon MCC every 10 steps,
    in the precision form with the Poisson solve by conjugate gradients
    preconditioned with FP32 V-cycles; `steps`
    steps timed by wall clock after `warmup` (electrons re-sorted every
    `resort` steps, the --resort-electrons of the cfg2 step),
    then as many with every phase between HIP events."""
    import torch
    from essaitemporaire_lepic_module_amd import config as cfg
    from essaitemporaire_lepic_module_amd.simulation import Simulation32
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats()
    t0 = time.perf_counter()
    case = cfg.cfg5_combined(n=n, np_cell=ppc)
    cells = cfg.loadable_cells(case.bcnd, case.n)
    sim = Simulation32(case, "cuda:0", capacity_factor=1.25,
                       resort_every=[resort, 0] if resort > 0 else None, poisson="pcg")
    torch.cuda.synchronize()
    t_init = time.perf_counter() - t0
    istep = 1
    sim.timed_phases = {"push_bc_deposit"}
    holes = torch.zeros_like(sim.lost_acc)
    for _ in range(warmup):
        sim.timers = {}
        holes += sim.lost_acc
        sim.advance_one_step(istep)
        istep += 1
    n_before = sim.particle_count()
    torch.cuda.synchronize()
    sim.timers = {}
    holes.zero_()
    slots = 0
    cycles = []
    t0 = time.perf_counter()
    for _ in range(steps):
        slots += sum(p.n for p in sim.parts)
        holes += sim.lost_acc
        sim.advance_one_step(istep)
        cycles.append(sim.last_mg["cycles"])
        istep += 1
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_phases = sim.phase_times()
    sim.timers = {}
    sim.timed_phases = None
    for _ in range(steps):
        sim.advance_one_step(istep)
        istep += 1
    torch.cuda.synchronize()
    phases = sim.phase_times()
    sim.check_escaped()
    push_ms, push_launches, push_units = kernel_phases.get("push_bc_deposit", (0.0, 0, 0))
    out = {"value": (slots - int(holes.sum().item())) / wall, "unit": "particle-steps/s",
           "ms_per_step": wall * 1e3 / steps, "steps": steps, "warmup": warmup,
           "ppc": ppc, "loadable_cells": cells, "particles": n_before,
           "electron_resort_every": resort,
           "pcg_iterations_per_solve": sum(cycles) / len(cycles),
           "phase_ms_per_step": {k: v[0] / steps for k, v in phases.items()},
           "push_GBs": (BYTES_PER_PARTICLE_STEP_F32 * push_units / (push_ms * 1e-3) / 1e9
                        if push_ms else 0.0),
           "init_s": t_init, "peak_device_GB": torch.cuda.max_memory_allocated() / 1e9,
           "workload": (f"{n}^3 cells, annular channel (r 0.3-0.45 of the width), anode 300 V at "
                        f"z = 0, radial Br(z) map (0.02 T at z = 0.7 L, width 0.1 L), {ppc} ppc "
                        f"x 2 species in the {cells} loadable cells, argon MCC every 10 steps, "
                        "FP32 particles, fixed-point deposit, Poisson to 1e-10 by FP64 conjugate "
                        "gradients preconditioned with one FP32 V-cycle")}
    del sim
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--ppc", type=int, default=64)
    ap.add_argument("--precision", choices=("f32", "f64", "both", "none"), default="both",
                    help="f32: the step's precision form (FP32 particles, fixed-point deposit, "
                         "FP64 Poisson) as the headline; f64: the FP64 step; both: f32 headline "
                         "with the FP64 step measured beside it; none: no cfg2 step, only the "
                         "other legs (for profiling them alone; the line then has no value)")
    ap.add_argument("--step-poisson", choices=("fp64", "pcg64"), default="fp64",
                    help="the precision-form step's Poisson solve: the reference's FP64 "
                         "multigrid, or FP64 conjugate gradients preconditioned with one FP64 "
                         "V-cycle of the same multigrid (FP64 throughout, same stopping rule)")
    ap.add_argument("--poisson-n", type=int, default=256,
                    help="grid of the mixed-precision Poisson comparison (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--combined-n", type=int, default=256,
                    help="grid of the combined run (annular channel, B map, mixed Poisson; "
                         "0: skip)")
    ap.add_argument("--combined-box-count", type=int, default=1,
                    help="also run the combined case with 32 ppc x 2 species of the whole box "
                         "(~1.07B particles at 256^3) placed in the channel (0: skip)")
    ap.add_argument("--combined-adjoint-n", type=int, default=64,
                    help="grid of the combined geometry's 500-step forward + adjoint of the "
                         "anode current (0: skip)")
    ap.add_argument("--adjoint-steps", type=int, default=200,
                    help="steps of the differentiable 64^3 run whose forward+adjoint / forward "
                         "cost is reported (0: skip)")
    ap.add_argument("--resort-electrons", type=int, default=4,
                    help="production path: also re-sort electrons by cell every k steps "
                         "(0: only at the reference's sort cadence)")
    ap.add_argument("--exact-order", action="store_true",
                    help="FP64 step: compact every step by look-back (the reference's particle "
                         "order at every step) instead of in place with holes removed at each sort")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the timed steps, copy 2 GiB device to device once: the "
                         "calibration kernel for rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    import torch
    from essaitemporaire_lepic_module_amd import config as cfg

    torch.cuda.init()
    if args.precision == "none":
        line = {"metric": "legs only (no cfg2 step)", "value": None}
        _other_legs(args, line)
        print(json.dumps(line))
        return
    head = "f64" if args.precision == "f64" else "f32"
    main_run = _step_bench(args, head)
    print(f"[bench] {head} step: {main_run['ms_per_step']:.2f} ms/step", file=sys.stderr,
          flush=True)
    other = None
    if args.precision == "both":
        other = _step_bench(args, "f64")
        print(f"[bench] f64 step: {other['ms_per_step']:.2f} ms/step", file=sys.stderr, flush=True)
    alt = None
    if args.precision == "both" and args.step_poisson == "fp64":
        # the headline's step with its Poisson solve by FP64 conjugate gradients
        alt_args = argparse.Namespace(**{**vars(args), "step_poisson": "pcg64"})
        alt = _step_bench(alt_args, "f32")
        print(f"[bench] f32 step, CG solve: {alt['ms_per_step']:.2f} ms/step", file=sys.stderr,
              flush=True)
    if args.copy_probe:
        a = torch.ones(COPY_PROBE_BYTES // 8, dtype=torch.float64, device="cuda:0")
        torch.empty_like(a).copy_(a)
        torch.cuda.synchronize()
        del a
    case = cfg.cfg2_pic_mcc(n=args.n, np_cell=args.ppc)
    solve = ("FP64 multigrid" if args.step_poisson == "fp64" else
             "FP64 conjugate gradients preconditioned with one FP64 V-cycle")
    prec = (f"FP32 particles (int32 cell + FP32 offsets), int64 fixed-point deposit, FP64 "
            f"Poisson ({solve}, to 1e-8)" if head == "f32" else "FP64 particles and fields")
    line = {
        "metric": "particle-steps/s of the 128^3 x 64 ppc x 2 species PIC-MCC step",
        "value": main_run["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_run["ms_per_step"],
        "step_ms_median": main_run["step_ms_median"],
        "step_ms_p90": main_run["step_ms_p90"],
        "step_ms_each": main_run["step_ms_each"],
        "first_timed_step": main_run["first_timed_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": head,
        "data": "synthetic (device Philox loader, Maxwellian e- 3 eV / Ar+ 0.03 eV)",
        "config": {"workload": "PIC-MCC at scale: 128^3 cells, 64 ppc x 2 species, Bz=0.01 T, "
                               f"Ar 1 Pa, MCC every 10, sort every 20, {prec}",
                   "cells": args.n ** 3, "particles": main_run["particles"], "ppc": args.ppc,
                   "electron_resort_every": main_run["electron_resort_every"],
                   "parallelism": "1 GPU"},
        "mg_solve_ms": main_run["mg_solve_ms"],
        "mg_vcycles_per_solve": main_run["mg_vcycles_per_solve"],
        "mg_roofline": main_run["mg_roofline"],
        "phase_ms_per_step": main_run["phase_ms_per_step"],
        "phase_timing": "second pass of the same number of steps, every phase between HIP "
                        "events (the timed pass brackets only the fused particle kernel)",
        "roofline": main_run["roofline"],
    }
    if other is not None:
        other["dtype"] = "f64"
        line["fp64_step"] = other
    if alt is not None:
        line["f32_step_cg_poisson"] = {
            k: alt[k] for k in ("value", "ms_per_step", "step_ms_median", "mg_solve_ms",
                                "mg_vcycles_per_solve", "phase_ms_per_step")}
        line["f32_step_cg_poisson"]["poisson"] = (
            "FP64 conjugate gradients preconditioned with one FP64 V-cycle, the same L1 "
            "residual criterion (1e-8); mg_vcycles_per_solve counts CG iterations")
    _other_legs(args, line)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(case)
    print(json.dumps(line))


def _other_legs(args, line):
    """The legs beside the cfg2 step: cfg3's Poisson solves, cfg5's combined
    run and its forward + adjoint, cfg4's forward + adjoint."""
    from essaitemporaire_lepic_module_amd import config as cfg
    if args.poisson_n:
        line["poisson_mixed"] = _poisson_cfg3(args.poisson_n)
        print("[bench] poisson done", file=sys.stderr, flush=True)
    if args.combined_n:
        line["combined"] = _combined_cfg5(args.combined_n, 32, resort=args.resort_electrons)
        print("[bench] combined done", file=sys.stderr, flush=True)
        if args.combined_box_count:
            # the same run with the particle count of 32 ppc x 2 species over
            # every cell of the box (the channel holds about a third of them)
            n = args.combined_n
            cells = cfg.loadable_cells(cfg.annular_channel_bcnd((n,) * 3, (3.125e-4,) * 3),
                                       (n,) * 3)
            line["combined_box_count"] = _combined_cfg5(n, int(round(32 * n ** 3 / cells)),
                                                        resort=args.resort_electrons)
            print("[bench] combined (box count) done", file=sys.stderr, flush=True)
    if args.adjoint_steps:
        line["adjoint"] = _adjoint_cfg4(args.adjoint_steps)
        # the same run with both Poisson solves by FP64 conjugate gradients
        cg = _adjoint_cfg4(args.adjoint_steps, poisson="pcg64")
        line["adjoint"]["cg_poisson"] = {
            k: cg[k] for k in ("forward_ms", "forward_checkpointed_ms", "backward_ms",
                               "adjoint_over_forward", "dL_dV", "dL_dB0")}
        print("[bench] adjoint done", file=sys.stderr, flush=True)
    if args.combined_adjoint_n:
        line["combined_adjoint"] = _adjoint_cfg5(args.combined_adjoint_n)
        print("[bench] combined adjoint done", file=sys.stderr, flush=True)


if __name__ == "__main__":
    main()
