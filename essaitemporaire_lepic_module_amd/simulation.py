"""Simulation%advance_one_step (modules/simulation/mod_simulation.f90:340-472)
on one GPU: every phase is a liblepic_hip.so launch on torch's current stream.

Step order (the modular driver, with the legacy Boris rotation in the
velocity update when B is on):
  sort (istep mod nb_step_sort == 1) -> reduce density -> rho -> Poisson ->
  E -> push -> wall BC -> MCC (istep mod nb_step_collisions == 0) -> deposit
  -> reduce density.
With fused=True the push, the wall BC and (on steps without MCC) the
deposit run as one kernel per species; the production path (lazy=True) also
deposits in that kernel on MCC steps, and MCC then adds the density of the
products it appends and takes out that of the particles it kills.  With lazy=True (production) that
kernel works in place and leaves holes where particles were lost; the
periodic sort removes them, keeping the survivors' order, and losses are
counted on the device without a host round trip.

resort_every (production only) re-sorts a species by cell every k steps in
addition to the reference's cadence: the deposit accumulates cell runs in
LDS, and fast species (electrons cross a cell in a few steps) lose their
cell order long before the reference's next sort.  The extra sorts change
the particle order between the reference's sorts, which nothing in the
production step depends on (collisions draw from per-particle Philox
streams and the deposit's summation order is that of its atomics anyway);
with them every sort is the counting sort by cell, whose order within a
cell is unspecified.
"""
from __future__ import annotations

import contextlib

import numpy as np

import torch

from . import _native
from . import fields as F
from . import particles as P
from .config import Case
from .device import as_device, node_shape


class Simulation:
    def __init__(self, case: Case, device=None, fused: bool = True, loader: str = "host",
                 mcc_mode: int = _native.MCC_REPLAY_RAN2, philox_key: int = 0x5EED5EED,
                 capacity_factor: float = 1.3, parts=None, np_thread=None, lazy: bool = False,
                 resort_every=None):
        self.case = case
        self.dev = torch.device(device if device is not None else "cuda:0")
        self.fused = fused
        self.lazy = lazy and fused
        if resort_every is not None and not self.lazy:
            raise ValueError("resort_every needs the in-place production path (lazy=True)")
        self.resort_every = list(resort_every) if resort_every is not None else None
        self.mcc_mode = mcc_mode
        self.philox_key = philox_key
        n, h = case.n, case.h
        self.n, self.h = n, h
        dev = self.dev
        self.bcnd = as_device(case.bcnd, torch.int32, dev)
        self.kq = F.build_kq(self.bcnd, n)
        self.phi = as_device(case.phi0(), torch.float64, dev)
        shape = node_shape(n)
        self.rho = torch.zeros(shape, dtype=torch.float64, device=dev)
        self.E4 = torch.zeros(shape + (4,), dtype=torch.float64, device=dev)
        tr = case.tracked
        self.ntype = len(tr)
        self.charge = [s.charge for s in tr]
        self.mass = [s.mass for s in tr]
        self.Nm = case.Nm[: self.ntype]
        self.np_red = torch.zeros((self.ntype,) + shape, dtype=torch.float64, device=dev)
        self.iseed = [P.seed_initialization(1, 0)]
        # room for np_cell particles in every cell the loader fills, with headroom
        from .config import loadable_cells
        cap = int(capacity_factor * abs(case.np_cell) * loadable_cells(case.bcnd, n)) + 1024
        if parts is not None:
            self.parts, self.np_thread = parts, np_thread
        elif loader == "host":
            self.parts, self.np_thread = P.load_particles_modular(
                n, h, case.bcnd, case.kq(), case.np_cell, case.vt0[: self.ntype], self.Nm,
                list(case.ni0) if case.ni0 is not None else [1.0] * self.ntype, self.iseed,
                x_load=case.x_load, cap=cap, device=dev)
        elif loader == "device":
            self.parts = P.load_uniform_device(n, h, self.bcnd, case.np_cell, case.vt0[: self.ntype],
                                               philox_key, cap, dev)
            self.np_thread = torch.zeros((self.ntype,) + shape, dtype=torch.float64, device=dev)
            for s, p in enumerate(self.parts):
                F.deposit_particle_set_to_np_thread(p, n, h, self.kq, self.Nm[s], self.np_thread[s])
        else:
            raise ValueError(f"unknown loader {loader!r}")
        F.reduce_species_density(n, self.bcnd, self.np_thread, self.ntype, self.np_red)
        self.p_mac = torch.zeros((self.ntype, case.ngrid + 1, 2), dtype=torch.float64, device=dev)
        self.P_loss = torch.zeros(self.ntype, dtype=torch.float64, device=dev)
        self.tables = None
        if case.chemistry is not None:
            self.Pcoll = torch.zeros(case.chemistry["ntype_total"], dtype=torch.float64, device=dev)
            self.tables = P.ReactionTables(case.chemistry, case.Nm, dev)
        self.tag_neg = -1
        for i, q in enumerate(self.charge):
            if q < 0.0 and i != 0:
                self.tag_neg = i + 1
                break
        self.bf = []
        for q in self.charge:
            if case.B_map is not None:
                self.B4 = as_device(case.B_map, torch.float64, dev)
                if case.B_grid is not None:
                    self.bf.append(P.bfield(3, B4=self.B4, magnetized=(q <= 0 or case.ions_magnetized),
                                            nB=case.B_grid[0], hB=case.B_grid[1]))
                else:
                    self.bf.append(P.bfield(2, B4=self.B4,
                                            magnetized=(q <= 0 or case.ions_magnetized)))
            elif case.B_uniform is not None:
                self.bf.append(P.bfield(1, case.B_uniform, magnetized=(q <= 0 or case.ions_magnetized)))
            else:
                self.bf.append(P.bfield(0))
        # dielectric surface charging: per-species charge planes the wall BC
        # fills and the next step's potential takes (mod_state.f90:148-157)
        self.dtype_d = self.sum_q_xz = self.sum_q_yz = None
        if case.flag_die:
            if case.dtype is None or len(case.dtype) != case.ngrid + 1:
                raise ValueError("flag_die needs the wall types of labels 0..ngrid")
            self.dtype_d = torch.as_tensor(np.asarray(case.dtype, np.int32), device=dev)
            n1, n2, n3 = (int(v) for v in n)
            self.sum_q_xz = torch.zeros((self.ntype, n3 + 3, n1 + 3), dtype=torch.float64, device=dev)
            self.sum_q_yz = torch.zeros((self.ntype, n3 + 3, n2 + 3, 2), dtype=torch.float64,
                                        device=dev)
        self.bp = [P.bc_params(n, h, case.flag_pbc, case.flag_nmn, s + 1, self.tag_neg,
                               self.Nm[s] * self.charge[s], self.mass[s], self.Nm[s], case.ngrid,
                               case.flag_die, self.dtype_d,
                               self.sum_q_xz[s] if case.flag_die else None,
                               self.sum_q_yz[s] if case.flag_die else None)
                   for s in range(self.ntype)]
        self.last_mg = None
        self.last_mcc = None
        self.lost = [0] * self.ntype
        # lazy mode: holes (lost slots not yet compacted) per species, on the device
        self.lost_acc = torch.zeros(self.ntype, dtype=torch.int64, device=dev)
        self.timers = None   # {phase: [(start_event, end_event, units), ...]} when timing
        self.timed_phases = None   # phases to time (None: all of them)
        self.diagnostics = None    # diagnostics.Diagnostics when attached
        # electron heating: absorbed energy (the reference's P_loss(2)) and the
        # R_ahp the window moments see, which the first heating sets
        self.P_heat = torch.zeros(1, dtype=torch.float64, device=dev)
        self.R_ahp_seen = 0.0
        self.last_heating = None

    @contextlib.contextmanager
    def _phase(self, name: str, units: int = 0):
        """Record HIP events around a phase on torch's current stream (the
        stream every lepic call is bound to)."""
        if not self._timing(name):
            yield
            return
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        yield
        b.record()
        self.timers.setdefault(name, []).append((a, b, units))

    def _timing(self, name: str) -> bool:
        return self.timers is not None and (self.timed_phases is None or name in self.timed_phases)

    def phase_times(self) -> dict:
        """{phase: (total_ms, launches, units)} of the recorded events."""
        torch.cuda.synchronize()
        out = {}
        for k, lst in (self.timers or {}).items():
            # (slots, holes tensor) pairs: device hole counts resolve here
            units = sum(u[0] - int(u[1]) if isinstance(u, tuple) else int(u) for _, _, u in lst)
            out[k] = (sum(a.elapsed_time(b) for a, b, _ in lst), len(lst), units)
        return out

    # ------------------------------------------------------------------
    def ensure_headroom(self):
        """Room for the products MCC appends: the larger of 2^20 slots and
        1/64 of the set, far above what a collision step creates at the
        reference's densities (the kernel reports ERR_CAPACITY rather than
        drop a product if it were ever exceeded)."""
        for p in self.parts:
            need = p.n + max(1 << 20, p.n // 64)
            if need > p.nmax:
                p.ensure_capacity(need)

    def advance_one_step(self, istep: int):
        c = self.case
        n, h = self.n, self.h
        if istep % c.nb_step_sort == 1:
            with self._phase("sort"):
                for s, p in enumerate(self.parts):
                    if self.lazy:
                        self._compact(s)
                    elif p.n > 1:
                        P.sort_particles_by_cell(p, n, h, cell_tables=False)
        elif self.resort_every is not None:
            with self._phase("sort"):
                for s in range(self.ntype):
                    k = self.resort_every[s]
                    if k and (istep - 1) % k == 0:
                        self._compact(s)
        with self._phase("rho"):
            F.reduce_species_density(n, self.bcnd, self.np_thread, self.ntype, self.np_red)
            F.build_rho_from_np(n, self.np_red, self.charge, self.ntype, self.rho)
        if c.flag_die:
            F.apply_dielectric_bc_to_phi(n, h, self.bcnd, self.dtype_d, c.ngrid, self.sum_q_xz,
                                         self.sum_q_yz, c.Ca * h[0] * h[2], self.phi)
        with self._phase("poisson"):
            self.last_mg = F.solve_poisson_legacy(self.phi, self.bcnd, self.rho, h, n, c.ncycl,
                                                  c.eps, c.omega, c.ng, c.flag_pbc, c.flag_nmn)
        with self._phase("efield"):
            F.calc_Efield_modular(n, h, self.phi, self.E4, self.bcnd)
        if self.diagnostics is not None:
            self.diagnostics.before_push(istep)
        if c.flag_die:
            self.sum_q_xz.zero_()
            self.sum_q_yz.zero_()
        collide = (self.tables is not None and c.nb_step_collisions > 0
                   and istep % c.nb_step_collisions == 0)
        # the production path deposits in the push on collision steps too and
        # lets MCC add its products to (and take its kills out of) that density
        dep_first = self.fused and self.lazy
        if self.fused:
            F.clear_np_thread(self.np_thread)
            push = P.push_bc_deposit_inplace if self.lazy else P.push_bc_deposit
            for s, p in enumerate(self.parts):
                np_s = None if collide and not dep_first else self.np_thread[s]
                if self.lazy:
                    # particles pushed = slots - holes before the call (read when timed)
                    phase = "push_bc_deposit" if np_s is not None else "push_bc"
                    units = (p.n, self.lost_acc[s].clone()) if self._timing(phase) else 0
                    with self._phase(phase, units):
                        P.push_bc_deposit_inplace(
                            p, n, h, self.E4, self.bf[s], self.charge[s], self.mass[s], c.dt,
                            self.bcnd, self.bp[s], self.kq, np_s, self.p_mac[s],
                            self.P_loss[s:s + 1], lost_acc=self.lost_acc[s:s + 1], sync=False)
                else:
                    with self._phase("push_bc_deposit" if not collide else "push_bc", p.n):
                        self.lost[s] = P.push_bc_deposit(
                            p, n, h, self.E4, self.bf[s], self.charge[s], self.mass[s], c.dt,
                            self.bcnd, self.bp[s], self.kq, np_s, self.p_mac[s],
                            self.P_loss[s:s + 1])
        else:
            with self._phase("push"):
                for s, p in enumerate(self.parts):
                    if self.bf[s].mode == 0:
                        P.move_particles_electrostatic(p, n, h, self.E4, self.charge[s],
                                                       self.mass[s], c.dt)
                    else:
                        P.move_particles_boris(p, n, h, self.E4, self.bf[s], self.charge[s],
                                               self.mass[s], c.dt)
            with self._phase("bc"):
                for s, p in enumerate(self.parts):
                    self.lost[s] = P.apply_particle_bc_legacy(
                        p, n, h, self.bcnd, c.flag_pbc, c.flag_nmn, s + 1, self.tag_neg,
                        c.flag_die, self.Nm[s] * self.charge[s], self.p_mac[s], self.mass[s],
                        self.P_loss[s:s + 1], self.Nm[s], self.dtype_d,
                        self.sum_q_xz[s] if c.flag_die else None,
                        self.sum_q_yz[s] if c.flag_die else None)
        if collide:
            self.ensure_headroom()
            with self._phase("mcc"):
                self.last_mcc = P.perform_collisions_step(
                    self.parts, n, h, self.np_red, self.tables, c.nb_step_collisions, c.dt,
                    self.iseed, self.Pcoll, mode=self.mcc_mode, philox_key=self.philox_key,
                    philox_counter=istep, np_dep=self.np_thread if dep_first else None,
                    kq=self.kq if dep_first else None)
        hc = c.heating
        if hc is not None and hc["nb_step_heating"] > 0 and istep % hc["nb_step_heating"] == 0:
            with self._phase("heating"):
                self._heat(istep)
        if (collide and not dep_first) or not self.fused:
            with self._phase("deposit"):
                F.clear_np_thread(self.np_thread)
                for s, p in enumerate(self.parts):
                    F.deposit_particle_set_to_np_thread(p, n, h, self.kq, self.Nm[s],
                                                        self.np_thread[s])
        with self._phase("rho"):
            F.reduce_species_density(n, self.bcnd, self.np_thread, self.ntype, self.np_red)
        if self.diagnostics is not None:
            self.diagnostics.after_step(istep)

    def _heat(self, istep: int):
        """compute_heating_region_moments, update_heating_vt and
        apply_electron_heating_local (mod_simulation.f90:441-447) on the
        electrons; velocities only, so the step's deposit stands."""
        hc = self.case.heating
        R_app = hc["R_ahp"] if hc["flag_circxh"] == 1 else self.R_ahp_seen
        hp = P.heating_params(hc["nudt"], hc["Pabs"], hc["nu_h"], hc["xl_pow"], hc["xr_pow"],
                              hc["ymax"], hc["zmax"], self.Nm[0], self.mass[0],
                              self.case.vt0[0], hc["flag_circxh"], hc["flag_ahp"],
                              self.R_ahp_seen, R_app)
        replay = self.mcc_mode == _native.MCC_REPLAY_RAN2
        self.last_heating = P.apply_electron_heating(
            self.parts[0], self.n, self.h, hp, self.iseed if replay else None, self.P_heat,
            self.mcc_mode, self.philox_key, istep, stats=replay)
        self.R_ahp_seen = R_app

    def run(self, nsteps: int, start: int = 1, out=None):
        """Steps start..start+nsteps-1 (mod_simulation.f90:475-494); with
        diagnostics attached, the summary of print_diagnostics is written to
        out (a text stream) before every step with mod(istep, nsav) == 1."""
        d = self.diagnostics
        for istep in range(start, start + nsteps):
            if d is not None and istep % d.nsav == 1:
                text = d.print_diagnostics(istep)
                if out is not None:
                    out.write(text)
            self.advance_one_step(istep)

    # ------------------------------------------------------------------
    def _compact(self, s: int):
        """Drop holes and re-sort by cell.  Without extra sorts the stable
        radix sort keeps the reference's particle order; once extra sorts are
        asked for, order within a cell is no longer the reference's anyway and
        every species is binned by the counting sort, which is cheaper."""
        p = self.parts[s]
        if p.n > 0:
            if self.resort_every is not None:
                P.bin_compact(p, self.n, self.h)
            else:
                P.sort_compact(p, self.n, self.h)
        self.lost_acc[s].zero_()

    @property
    def holes(self) -> list:
        """Lost slots not yet compacted, per species (lazy mode; reads the device)."""
        return [int(v) for v in self.lost_acc.tolist()]

    def particle_count(self) -> int:
        """Live particles (slots minus holes not yet compacted)."""
        return sum(p.n for p in self.parts) - int(self.lost_acc.sum().item())

    def total_charge(self) -> float:
        """sum_s q_s Nm_s N_s."""
        return float(sum(q * nm * p.n for q, nm, p in zip(self.charge, self.Nm, self.parts)))


_SIDE = {}


def _side_worker(dev):
    """The process's one worker thread for side-stream work (it gets its own
    lepic context) and the side stream of `dev` (default priority)."""
    import concurrent.futures
    key = torch.device(dev).index or 0
    if key not in _SIDE:
        _SIDE[key] = (concurrent.futures.ThreadPoolExecutor(1), torch.cuda.Stream(dev))
    return _SIDE[key]


class Simulation32(Simulation):
    """The step in its precision form: particles as int32 cell + FP32 in-cell
    offsets + FP32 velocities (particles32.ParticleSet32), the FP32 gather +
    push + wall BC, the 64-bit fixed-point deposit (density independent of
    particle order, charge exactly conserved), FP64 fields.  Collisions run
    in PHILOX mode on the FP32 sets and adjust the fixed-point density for
    the products they append and the particles they kill.  The order of the
    step is Simulation's production path; every sort is the stable sort by
    cell, so a run is reproducible bit for bit.

    poisson: "fp64" (the reference's multigrid), "mixed" (FP32 smoothing
    inside FP64 iterative refinement, fields.solve_poisson_mixed) or "pcg"
    (FP64 conjugate gradients preconditioned with the FP32 V-cycles,
    fields.solve_poisson_pcg; plasma and electrode labels only) or "pcg64"
    (the same with FP64 V-cycles: FP64 throughout)."""

    def __init__(self, case: Case, device=None, philox_key: int = 0x5EED5EED,
                 capacity_factor: float = 1.3, resort_every=None, poisson: str = "fp64",
                 parts64=None, overlap_sort: bool = True):
        from . import particles32 as P32
        if case.flag_die or case.heating is not None:
            raise ValueError("the FP32 step has no dielectric charging or electron heating")
        if case.B_grid is not None:
            raise ValueError("the FP32 step takes B uniform or as a map on the field grid")
        super().__init__(case, device, fused=True, loader="device" if parts64 is None else "host",
                         mcc_mode=_native.MCC_PHILOX, philox_key=philox_key,
                         capacity_factor=capacity_factor, lazy=True, resort_every=None,
                         parts=parts64, np_thread=None if parts64 is None else torch.zeros(
                             (len(case.tracked),) + node_shape(case.n), dtype=torch.float64,
                             device=device if device is not None else "cuda:0"))
        if poisson not in ("fp64", "mixed", "pcg", "pcg64"):
            raise ValueError(f"unknown Poisson solve {poisson!r}")
        self.poisson = poisson
        # the step's sorts touch only the particles and the field phase only
        # the grid: with overlap_sort the sorts run on a side stream, driven
        # by a worker thread with its own lepic context, while this thread
        # drives the density, the solve and E (unless the sort phase is timed)
        self.overlap_sort = overlap_sort
        self.resort_every32 = list(resort_every) if resort_every is not None else None
        n, h = self.n, self.h
        shape = node_shape(n)
        dev = self.dev
        self.frac_bits = [F.fixed_point_bits32(n, h, nm) for nm in self.Nm]
        parts32 = []
        for p in self.parts:
            q, out = P32.ParticleSet32.from_pset64(p, n, h, cap=max(p.nmax, 16))
            if out:
                raise ValueError(f"{out} particles lie outside the cells")
            if parts64 is None:
                p.release()                # the FP64 load is only the source
            P32.sort32(q, n, h)
            parts32.append(q)
        self.parts = parts32
        self.dep = torch.zeros((self.ntype,) + shape, dtype=torch.int64, device=dev)
        for s, p in enumerate(self.parts):
            P32.deposit32(p, n, h, self.Nm[s], self.frac_bits[s], self.dep[s])
        self.np_thread.zero_()
        for s in range(self.ntype):
            P32.fixed_density(n, h, self.dep[s], self.kq, self.frac_bits[s], self.np_thread[s])
        F.reduce_species_density(n, self.bcnd, self.np_thread, self.ntype, self.np_red)
        self.E4f = torch.zeros(shape + (4,), dtype=torch.float32, device=dev)
        self.escaped = torch.zeros(self.ntype, dtype=torch.int64, device=dev)
        self.B4f = P32.round_field4(self.B4) if case.B_map is not None else None
        self.bf32 = []
        for q in self.charge:
            mag = q <= 0 or case.ions_magnetized
            if case.B_map is not None:
                self.bf32.append(P32.bfield32(2, B4f=self.B4f, magnetized=mag))
            elif case.B_uniform is not None:
                self.bf32.append(P32.bfield32(1, case.B_uniform, magnetized=mag))
            else:
                self.bf32.append(P32.bfield32(0))
        self._P32 = P32

    def ensure_headroom(self):
        for p in self.parts:
            need = p.n + max(1 << 20, p.n // 64)
            if need > p.nmax:
                p.ensure_capacity(need)

    def _compact(self, s: int):
        p = self.parts[s]
        if p.n > 0:
            self._P32.sort32(p, self.n, self.h)
        self.lost_acc[s].zero_()

    def check_escaped(self):
        """Raises when a survivor of a push left the cells (a domain without
        walls around it); reads the device."""
        esc = int(self.escaped.sum().item())
        if esc:
            raise RuntimeError(f"{esc} particles left the domain without meeting a wall")

    def _sorts_due(self, istep: int) -> list:
        if istep % self.case.nb_step_sort == 1:
            return list(range(self.ntype))
        if self.resort_every32 is None:
            return []
        return [s for s in range(self.ntype)
                if self.resort_every32[s] and (istep - 1) % self.resort_every32[s] == 0]

    def _start_sorts(self, due: list):
        """The sorts of `due` on the side stream, after the work already on
        this thread's stream; returns the future of their completion event."""
        pool, side = _side_worker(self.dev)
        ready = torch.cuda.Event()
        ready.record(torch.cuda.current_stream(self.dev))
        dev = self.dev

        def work():
            with torch.cuda.device(dev), torch.cuda.stream(side):
                side.wait_event(ready)
                for s in due:
                    self._compact(s)
                done = torch.cuda.Event()
                done.record(side)
            return done
        return pool.submit(work)

    def advance_one_step(self, istep: int):
        c = self.case
        n, h = self.n, self.h
        P32 = self._P32
        due = self._sorts_due(istep)
        sorting = None
        if due and self.overlap_sort and not self._timing("sort"):
            sorting = self._start_sorts(due)
        elif due:
            with self._phase("sort"):
                for s in due:
                    self._compact(s)
        with self._phase("rho"):
            F.reduce_species_density(n, self.bcnd, self.np_thread, self.ntype, self.np_red)
            F.build_rho_from_np(n, self.np_red, self.charge, self.ntype, self.rho)
        with self._phase("poisson"):
            if self.poisson == "mixed":
                self.last_mg = F.solve_poisson_mixed(self.phi, self.bcnd, self.rho, h, n, c.ncycl,
                                                     c.eps, c.omega, c.ng, c.flag_pbc, c.flag_nmn)
            elif self.poisson in ("pcg", "pcg64"):
                self.last_mg = F.solve_poisson_pcg(self.phi, self.bcnd, self.rho, h, n, c.ncycl,
                                                   c.eps, c.omega, c.ng, c.flag_pbc, c.flag_nmn,
                                                   precond_fp64=self.poisson == "pcg64")
            else:
                self.last_mg = F.solve_poisson_legacy(self.phi, self.bcnd, self.rho, h, n, c.ncycl,
                                                      c.eps, c.omega, c.ng, c.flag_pbc, c.flag_nmn)
        with self._phase("efield"):
            F.calc_Efield_modular(n, h, self.phi, self.E4, self.bcnd)
            P32.round_field4(self.E4, self.E4f)
        collide = (self.tables is not None and c.nb_step_collisions > 0
                   and istep % c.nb_step_collisions == 0)
        self.dep.zero_()
        # the species being sorted push last, so that their sorts also
        # overlap the other species' pushes
        order = list(range(self.ntype))
        if sorting is not None:
            order = [s for s in order if s not in due] + [s for s in order if s in due]
        for s in order:
            if sorting is not None and s in due:
                torch.cuda.current_stream(self.dev).wait_event(sorting.result())
                sorting = None
            p = self.parts[s]
            units = (p.n, self.lost_acc[s].clone()) if self._timing("push_bc_deposit") else 0
            with self._phase("push_bc_deposit", units):
                P32.push32(p, n, h, self.E4f, self.bf32[s], self.charge[s], self.mass[s], c.dt,
                           self.bcnd, self.bp[s], self.frac_bits[s], self.dep[s], self.p_mac[s],
                           self.P_loss[s:s + 1], self.lost_acc[s:s + 1],
                           self.escaped[s:s + 1])
        if collide:
            self.ensure_headroom()
            with self._phase("mcc"):
                self.last_mcc = P32.collisions32(self.parts, n, h, self.np_red, self.tables,
                                                 c.nb_step_collisions, c.dt, self.Pcoll,
                                                 self.philox_key, istep, self.dep, self.frac_bits)
        with self._phase("rho"):
            self.np_thread.zero_()
            for s in range(self.ntype):
                P32.fixed_density(n, h, self.dep[s], self.kq, self.frac_bits[s], self.np_thread[s])
            F.reduce_species_density(n, self.bcnd, self.np_thread, self.ntype, self.np_red)
