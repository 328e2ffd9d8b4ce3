"""Differentiation of the PIC step: dL/dtheta of a wall-flux loss with
respect to an electrode's bias V and the uniform field magnitude B0, by
hand-written adjoint kernels (lepic_adj_*) run backward over the steps.

The forward is the FP64 step on particles held in place (lost particles
become holes and keep their slots, so each slot is one particle for the
whole run): rho from the deposit, the multigrid potential with the biased
electrode at V, E, the Boris push with B = (0, 0, B0) (or the case's B map,
gathered like E) and the E gather, the
wall test, the fused deposit for the next step, and on collision steps the
null-collision MCC (mod_simulation.f90:340-472 without sorting; MCC
products take slots reserved after the initial particles).  The loss is the
time-averaged flux of one species (Ar+) through the node plane one cell in
front of the biased wall,

    L = (1/K) sum_k q Nm / (Ly Lz h1) sum_p vx_p hat(x_p / h1 - (n1 - 1)),

or, with the case's flux_plane (axis a, node plane ip, sign, area A), the
same sum of sign q Nm / (A h_a) v_a hat(x_a / h_a - (ip - 1)) -- with A = 1
the ion current into an electrode face (the annular channel's anode);
a smooth stand-in for the charge the wall collects per unit time and area
(apply_particle_bc_legacy's p_mac, mod_particleBC.f90:195-206, counts whole
particles and has zero derivative almost everywhere).  The events are
frozen in the backward: which particles the walls absorb and at which step,
and the collisions -- which particles collide at which step, the velocity
each leaves with and the products (their whole state) -- are constants of
the map the gradient differentiates.  So a collided particle's velocity
carries no derivative from before its collision (its position does), and a
product's state none from before its creation.

Memory: particle states are kept only at checkpoints every `checkpoint`
steps (positions, velocities, hole flags, the deposited density and the
potential); the backward recomputes each segment forward from its
checkpoint, keeping that segment's states, then walks it backward.

PicLoss is the torch.autograd.Function: forward(V, B0) -> L, backward
-> (dL/dV, dL/dB0)."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _native
from . import fields as F
from . import particles as P
from .config import EPS0, Case, node_shape
from .device import context, ptr

_STATE = ("x", "y", "z", "vx", "vy", "vz", "dead")


def _pset_of(d: dict, n: int, cap: int | None = None) -> _native.PSet:
    """The set of the first n slots of the arrays in d (cap slots)."""
    p = _native.PSet()
    for f in ("x", "y", "z", "vx", "vy", "vz", "dead"):
        setattr(p, f, d[f].data_ptr())
    p.cell = None
    p.n = n
    p.cap = n if cap is None else cap
    return p


class AdjointSet:
    """dL/dx and dL/dv per particle slot of one species (device FP64)."""

    def __init__(self, n: int, device):
        self.n = n
        self.a = {k: torch.zeros(max(n, 1), dtype=torch.float64, device=device)
                  for k in ("lx", "ly", "lz", "lvx", "lvy", "lvz")}

    def zero_(self):
        for v in self.a.values():
            v.zero_()

    def unpermute_(self, perm: torch.Tensor):
        """Slot perm[j] takes slot j's values (j < len(perm)); every other
        slot is zero."""
        m = perm.numel()
        for k, v in self.a.items():
            out = torch.zeros_like(v)
            out.index_copy_(0, perm, v[:m])
            self.a[k] = out

    def struct(self) -> _native.AdjSet:
        s = _native.AdjSet()
        for k, v in self.a.items():
            setattr(s, k, v.data_ptr())
        s.n = self.n
        return s


class DiffPIC:
    """The differentiable run of a Case on one device.

    case: plasma (-1) and electrode (>= 1) labels only, B uniform along z, a
    B map on the field grid (B_map; dL/dB0 is then 0) or none, collisions (PHILOX draws, keyed by philox_key and the step) or
    none, no dielectrics or heating.  poisson: "fp64" (the reference's
    multigrid), "pcg" (conjugate gradients preconditioned with FP32
    V-cycles, to the same tolerance) or "pcg64" (the same with FP64
    V-cycles) for the forward's solves (the backward's recomputed
    segments use the same); adjoint_poisson: the same choice for the
    adjoint solves A mu = dL/dphi, which have no counterpart in the
    reference (default: FP64 conjugate gradients preconditioned with FP32
    V-cycles).
    resort: every `resort` steps (0: never) the live particles are sorted by
    cell at the start of the step and the holes dropped, so that gathers,
    deposits and the adjoint's scatter stay local; the slot orders are
    recorded with the run's events, replayed by frozen runs and undone on
    the adjoint in the backward.
    parts: the initial FP64 particle sets
    (host arrays dict per species: x y z vx vy vz), loaded into device sets
    of that many slots plus, with collisions, room for the products
    (product_room of the initial count, at least 1024 slots)."""

    def __init__(self, case: Case, parts, device=None, bias_label: int = 2, flux_species: int = 1,
                 checkpoint: int = 20, adjoint_eps: float = 1e-10, philox_key: int = 0x5EED5EED,
                 product_room: float = 1.0 / 16, poisson: str = "fp64", resort: int = 20,
                 adjoint_poisson: str = "pcg"):
        for sv in (poisson, adjoint_poisson):
            if sv not in ("fp64", "pcg", "pcg64"):
                raise ValueError(f"unknown Poisson solve {sv!r}")
        self.poisson = poisson
        self.adjoint_poisson = adjoint_poisson
        self.R = max(0, int(resort))
        self._track = True
        b = np.asarray(case.bcnd)
        if np.any((b == 0) | (b == -2)) or case.flag_pbc or case.flag_nmn:
            raise ValueError("the adjoint E field takes plasma (-1) and electrode labels only")
        if case.flag_die or case.heating is not None:
            raise ValueError("the differentiable step has no dielectrics or heating")
        if case.B_map is not None and case.B_grid is not None:
            raise ValueError("the differentiable step takes a B map on the field grid")
        if case.B_uniform is not None and (case.B_uniform[0] or case.B_uniform[1]):
            raise ValueError("the differentiable step takes B along z")
        self.case = case
        self.dev = torch.device(device if device is not None else "cuda:0")
        self.ctx = context(self.dev)
        self.n, self.h = case.n, case.h
        self.g = _native.grid(self.n, self.h)
        self.label = int(bias_label)
        self.flux_species = int(flux_species)
        self.K = max(1, int(checkpoint))
        self.adj_eps = float(adjoint_eps)
        dev = self.dev
        self.bcnd = torch.as_tensor(b.astype(np.int32), device=dev)
        self.kq = torch.as_tensor(case.kq(), device=dev)
        self.bias_mask = self.bcnd == self.label
        self.elec_mask = torch.zeros_like(self.bcnd, dtype=torch.bool)
        for lab in case.electrode_V:
            self.elec_mask |= self.bcnd == lab
        tr = case.tracked
        self.ntype = len(tr)
        self.charge = [s.charge for s in tr]
        self.mass = [s.mass for s in tr]
        self.Nm = list(case.Nm[: self.ntype])
        self.magnetized = [q <= 0 or case.ions_magnetized for q in self.charge]
        # a B map (on the field grid) replaces the uniform B0 along z; dL/dB0
        # is then 0
        self.B4 = None
        if case.B_map is not None:
            self.B4 = torch.as_tensor(np.asarray(case.B_map, np.float64), device=dev).contiguous()
        shape = node_shape(self.n)
        self.shape = shape
        self.collide_every = int(case.nb_step_collisions) if case.chemistry is not None else 0
        self.tables = None
        if self.collide_every > 0:
            self.tables = P.ReactionTables(case.chemistry, case.Nm, dev)
            self.Pcoll = torch.zeros(case.chemistry["ntype_total"], dtype=torch.float64,
                                     device=dev)
        self.philox_key = int(philox_key)
        self.init = []
        # slots hold the particles sorted by cell (the loss and its gradient
        # do not depend on the order; gathers and deposits of cell-sorted
        # particles reuse stencils); perm[s][slot] is the input index
        self.perm = []
        n1, n2, n3 = case.n
        for a in parts:
            m = len(a["x"])
            cx = np.clip((np.asarray(a["x"]) / case.h[0]).astype(np.int64), 0, n1 - 1)
            cy = np.clip((np.asarray(a["y"]) / case.h[1]).astype(np.int64), 0, n2 - 1)
            cz = np.clip((np.asarray(a["z"]) / case.h[2]).astype(np.int64), 0, n3 - 1)
            order = np.argsort(cx + n1 * (cy + n2 * cz), kind="stable")
            self.perm.append(order)
            a = {f: np.asarray(a[f], np.float64)[order] for f in ("x", "y", "z", "vx", "vy", "vz")}
            cap = max(m + (max(1024, int(m * product_room)) if self.collide_every else 0), 4)
            d = {}
            for f in ("x", "y", "z", "vx", "vy", "vz"):
                d[f] = torch.zeros(cap, dtype=torch.float64, device=dev)
                d[f][:m] = torch.as_tensor(np.asarray(a[f], np.float64), device=dev)
            d["dead"] = torch.ones(cap, dtype=torch.int8, device=dev)
            d["dead"][:m] = 0
            self.init.append((m, d, cap))
        self.mcc_events = {}
        n1, n2, n3 = self.n
        # the loss plane: the case's (axis, node plane, sign, area), else the
        # node plane x = (n1 - 1) h1 over the x = L face's area
        axis, plane, sign, area = case.flux_plane or (0, n1, 1.0, None)
        if area is None:
            area = math.prod(self.n[d] * self.h[d] for d in range(3) if d != axis)
        self.flux_axis, self.plane = int(axis), int(plane)
        self.flux_coef = float(sign) * self.charge[self.flux_species] * \
            self.Nm[self.flux_species] / (area * self.h[axis])
        self.phi = torch.zeros(shape, dtype=torch.float64, device=dev)
        self.rho = torch.zeros(shape, dtype=torch.float64, device=dev)
        self.E4 = torch.zeros(shape + (4,), dtype=torch.float64, device=dev)
        self.np_thread = torch.zeros((self.ntype,) + shape, dtype=torch.float64, device=dev)
        self.np_red = torch.zeros_like(self.np_thread)
        self.p_mac = torch.zeros((self.ntype, 2 * (case.ngrid + 1)), dtype=torch.float64,
                                 device=dev)
        self.P_loss = torch.zeros(self.ntype, dtype=torch.float64, device=dev)
        self.lost = torch.zeros(self.ntype, dtype=torch.int64, device=dev)
        self.acc = torch.zeros(1, dtype=torch.float64, device=dev)
        self.timers = None

    # -- forward ----------------------------------------------------------
    def _fresh(self):
        """Working particle arrays from the initial sets ([slots in use,
        arrays] per species) and the initial deposit."""
        st = []
        self.ident, self.next_id = [], []
        for m, d, cap in self.init:
            w = {k: v.clone() for k, v in d.items()}
            st.append([m, w])
            # the particle each slot holds: its input index, products
            # numbered from m on in order of creation
            self.ident.append(torch.arange(cap, dtype=torch.int64, device=self.dev))
            self.next_id.append(m)
        self.np_thread.zero_()
        for s, (m, w) in enumerate(st):
            p = _pset_of(w, m)
            self.ctx.check(self.ctx.lib.lepic_deposit(self.ctx.h, C.byref(self.g), C.byref(p),
                                                      ptr(self.kq), float(self.Nm[s]),
                                                      ptr(self.np_thread[s])))
        return st

    def by_input(self, s: int, per_slot) -> np.ndarray:
        """A per-slot array of species s (slots 0..m-1, host or device) in the
        order of the particles the run was given."""
        v = per_slot.cpu().numpy() if isinstance(per_slot, torch.Tensor) else np.asarray(per_slot)
        order = self.perm[s]
        out = np.empty((len(order),) + v.shape[1:], v.dtype)
        out[order] = v[:len(order)]
        return out

    def _set_electrodes(self, phi, V):
        """Electrode nodes take their voltages (the biased label V), without
        a host synchronisation."""
        if getattr(self, "_elec_V", None) != float(V):
            vmap = torch.zeros_like(phi)
            for lab, v in self.case.electrode_V.items():
                vmap.masked_fill_(self.bcnd == lab, float(V) if lab == self.label else float(v))
            self._elec_map, self._elec_V = vmap, float(V)
        torch.where(self.elec_mask, self._elec_map, phi, out=phi)

    def _fields(self, V):
        """rho from the deposited density, phi (warm start), E."""
        c = self.case
        F.reduce_species_density(self.n, self.bcnd, self.np_thread, self.ntype, self.np_red)
        F.build_rho_from_np(self.n, self.np_red, self.charge, self.ntype, self.rho)
        self._set_electrodes(self.phi, V)
        self._solve(self.phi, self.rho, c.eps, self.poisson)
        F.calc_Efield_modular(self.n, self.h, self.phi, self.E4, self.bcnd)

    def _solve(self, u, rhs, eps, how):
        """A u = rhs on plasma nodes to eps (warm start from u): the FP64
        multigrid ("fp64"), or the conjugate gradients preconditioned with
        FP32 ("pcg") or FP64 ("pcg64") V-cycles."""
        c = self.case
        if how in ("pcg", "pcg64"):
            F.solve_poisson_pcg(u, self.bcnd, rhs, self.h, self.n, c.ncycl, eps, c.omega, c.ng,
                                precond_fp64=how == "pcg64")
        else:
            F.solve_poisson_legacy(u, self.bcnd, rhs, self.h, self.n, c.ncycl, eps, c.omega,
                                   c.ng, 0, 0)

    def _bf(self, s: int, B0):
        """The push's B for species s: the case's map, else B0 along z (None:
        no B)."""
        if self.B4 is not None:
            return P.bfield(2, B4=self.B4, magnetized=self.magnetized[s])
        if B0 is None:
            return P.bfield(0)
        return P.bfield(1, (0.0, 0.0, float(B0)), magnetized=self.magnetized[s])

    def _push(self, st, B0):
        c = self.case
        self.np_thread.zero_()
        for s, (m, w) in enumerate(st):
            bf = self._bf(s, B0)
            p = _pset_of(w, m)
            bp = P.bc_params(self.n, self.h, 0, 0, s + 1, -1, self.Nm[s] * self.charge[s],
                             self.mass[s], self.Nm[s], c.ngrid)
            self.ctx.check(self.ctx.lib.lepic_push_bc_deposit_inplace(
                self.ctx.h, C.byref(self.g), C.byref(p), ptr(self.E4), C.byref(bf),
                self.charge[s], self.mass[s], c.dt, ptr(self.bcnd), C.byref(bp), ptr(self.kq),
                ptr(self.np_thread[s]), ptr(self.p_mac[s]), ptr(self.P_loss[s:s + 1]),
                ptr(self.lost[s:s + 1]), None))

    def _flux(self, st, acc, adj=None, scale=0.0):
        m, w = st[self.flux_species]
        p = _pset_of(w, m)
        self.ctx.check(self.ctx.lib.lepic_plane_flux_axis(
            self.ctx.h, C.byref(self.g), C.byref(p), self.flux_axis, self.plane,
            float(self.flux_coef),
            ptr(acc), C.byref(adj.struct()) if adj is not None else None, float(scale)))

    def _sorts_at(self, k: int) -> bool:
        return self.R > 0 and k > 1 and (k - 1) % self.R == 0

    def _sort_perms(self, st):
        """Per species, the live slots in cell order (stable): the new
        order's old slot numbers."""
        n1, n2, n3 = self.n
        perms = []
        for m, w in st:
            live = (w["dead"][:m] == 0).nonzero().squeeze(1)
            cx = torch.clamp((w["x"][live] / self.h[0]).floor().long(), 0, n1 - 1)
            cy = torch.clamp((w["y"][live] / self.h[1]).floor().long(), 0, n2 - 1)
            cz = torch.clamp((w["z"][live] / self.h[2]).floor().long(), 0, n3 - 1)
            order = torch.argsort(cx + n1 * (cy + n2 * cz), stable=True)
            perms.append(live[order])
        return perms

    def _apply_sort(self, st, perms):
        """Slots j < len(perm) take the particle of slot perm[j]; the rest
        are unused."""
        for s, perm in enumerate(perms):
            m, w = st[s]
            m1 = perm.numel()
            for f in ("x", "y", "z", "vx", "vy", "vz"):
                w[f][:m1] = w[f].index_select(0, perm)
            w["dead"][:m1] = 0
            w["dead"][m1:m] = 1
            if self._track:
                self.ident[s][:m1] = self.ident[s].index_select(0, perm)
            st[s][0] = m1

    def _new_products(self, s: int, m0: int, m1: int):
        if self._track and m1 > m0:
            self.ident[s][m0:m1] = torch.arange(self.next_id[s], self.next_id[s] + m1 - m0,
                                                dtype=torch.int64, device=self.dev)
            self.next_id[s] += m1 - m0

    @staticmethod
    def _snap(st):
        return [[m, {k: v.clone() for k, v in w.items()}] for m, w in st]

    def _collides(self, k: int) -> bool:
        return self.collide_every > 0 and k % self.collide_every == 0

    def _collide(self, st, k: int, events=None):
        """MCC at step k (perform_collisions_step, mod_collisions.f90:214-739,
        on all particles with per-particle Philox draws), its density change
        added to the step's deposit.  events None: run it and return the
        record (per species: the collided slots and the velocities they
        leave with, the product slots and their states); else replay that
        record (the frozen-event run)."""
        ctx, g = self.ctx, self.g
        if events is not None:
            for s, ev in enumerate(events):
                m, w = st[s]
                if ev["idx"].numel():
                    for f in ("vx", "vy", "vz"):
                        w[f].index_copy_(0, ev["idx"], ev[f])
                a, b = ev["m0"], ev["m1"]
                if b > a:
                    for f in ("x", "y", "z", "vx", "vy", "vz"):
                        w[f][a:b] = ev["prod"][f]
                    w["dead"][a:b] = 0
                    # the record's own (aligned) arrays: the deposit loads 16-byte vectors
                    p = _pset_of(ev["prod"], b - a)
                    ctx.check(ctx.lib.lepic_deposit(ctx.h, C.byref(g), C.byref(p), ptr(self.kq),
                                                    float(self.Nm[s]), ptr(self.np_thread[s])))
                self._new_products(s, a, b)
                st[s][0] = b
            return events
        c = self.case
        v0 = [torch.stack([w["vx"][:m], w["vy"][:m], w["vz"][:m]]) for m, w in st]
        dead0 = [w["dead"][:m].clone() for m, w in st]
        arr = (_native.PSet * self.ntype)(*[_pset_of(w, m, w["x"].numel()) for m, w in st])
        seed = C.c_int32(0)
        stats = _native.MCCStats()
        ctx.check(ctx.lib.lepic_collisions_deposited(
            ctx.h, C.byref(g), arr, ptr(self.np_red), C.byref(self.tables.c),
            int(c.nb_step_collisions), float(c.dt), int(_native.MCC_PHILOX), C.byref(seed),
            C.c_uint64(self.philox_key), C.c_uint64(k), ptr(self.Pcoll), C.byref(stats),
            ptr(self.np_thread), ptr(self.kq)))
        out = []
        for s, (m, w) in enumerate(st):
            m1 = int(arr[s].n)
            if not torch.equal(w["dead"][:m], dead0[s]):
                raise RuntimeError("collisions removed particles; the differentiable step "
                                   "freezes velocity changes and products only")
            v1 = torch.stack([w["vx"][:m], w["vy"][:m], w["vz"][:m]])
            idx = (v1 != v0[s]).any(0).nonzero().squeeze(1)
            prod = {f: w[f][m:m1].clone() for f in ("x", "y", "z", "vx", "vy", "vz")}
            prod["dead"] = torch.zeros(m1 - m + 16, dtype=torch.int8, device=self.dev)
            ev = {"idx": idx, "m0": m, "m1": m1, "prod": prod}
            for f in ("vx", "vy", "vz"):
                ev[f] = w[f].index_select(0, idx)
            out.append(ev)
            self._new_products(s, m, m1)
            st[s][0] = m1
        return out

    def _push_frozen(self, st, B0, deaths, k):
        """The push with the wall events of a recorded run: every live slot
        moves, the slots recorded lost at step k become holes, the rest
        deposit (the map the adjoint differentiates)."""
        c = self.case
        self.np_thread.zero_()
        for s, (m, w) in enumerate(st):
            bf = self._bf(s, B0)
            p = _pset_of(w, m)
            self.ctx.check(self.ctx.lib.lepic_move_boris(self.ctx.h, C.byref(self.g), C.byref(p),
                                                         ptr(self.E4), C.byref(bf), self.charge[s],
                                                         self.mass[s], c.dt))
            per_slot = deaths[s].index_select(0, self.ident[s][:m]).contiguous()
            self.ctx.check(self.ctx.lib.lepic_apply_deaths(self.ctx.h, C.byref(p),
                                                           ptr(per_slot), int(k)))
            self.ctx.check(self.ctx.lib.lepic_deposit(self.ctx.h, C.byref(self.g), C.byref(p),
                                                      ptr(self.kq), float(self.Nm[s]),
                                                      ptr(self.np_thread[s])))

    def forward(self, V: float, B0, steps: int, keep: bool = True,
                record_deaths: bool = False, frozen=None, frozen_mcc=None,
                frozen_sorts=None) -> float:
        """The loss of `steps` steps at bias V and field B0 (None: no B);
        keep: store the checkpoints the backward needs; record_deaths: keep
        in self.deaths, per species, the step at which each particle was
        lost (0: never), indexed by particle (input index, then products in
        order of creation); frozen: such a record, whose wall events replace
        the wall test; frozen_mcc: a run's self.mcc_events, whose collisions
        replace MCC, and frozen_sorts its self.sorts (default: self.sorts),
        whose slot orders replace the sorts (all three: the run the adjoint
        differentiates, for checks by finite differences).  Without
        frozen_mcc the collisions are drawn and recorded in self.mcc_events
        and the sorts in self.sorts."""
        self._track = True
        st = self._fresh()
        self.phi.zero_()
        self.acc.zero_()
        self.ckpt = {}
        if frozen_mcc is None:
            self.mcc_events = {}
            self.sorts = {}
            sorts = None
        else:
            sorts = frozen_sorts if frozen_sorts is not None else self.sorts
        if record_deaths:
            self.deaths = [torch.zeros(max(2 * w["x"].numel(), 1), dtype=torch.int64,
                                       device=self.dev) for _, w in st]
        for k in range(1, steps + 1):
            if self._sorts_at(k):
                if sorts is None:
                    self.sorts[k] = self._sort_perms(st)
                    self._apply_sort(st, self.sorts[k])
                else:
                    self._apply_sort(st, sorts[k])
            if keep and (k - 1) % self.K == 0:
                self.ckpt[k] = (self._snap(st), self.np_thread.clone(), self.phi.clone())
            self._fields(V)
            if frozen is not None:
                self._push_frozen(st, B0, frozen, k)
            else:
                self._push(st, B0)
            if self._collides(k):
                if frozen_mcc is not None:
                    self._collide(st, k, frozen_mcc[k])
                else:
                    self.mcc_events[k] = self._collide(st, k)
            self._flux(st, self.acc)
            if record_deaths:
                for s, (m, w) in enumerate(st):
                    if self.next_id[s] > self.deaths[s].numel():
                        self.deaths[s] = torch.cat([self.deaths[s],
                                                    torch.zeros_like(self.deaths[s])])
                    ids = self.ident[s][:m]
                    d = self.deaths[s].index_select(0, ids)
                    lost = (w["dead"][:m] != 0) & (d == 0)
                    self.deaths[s].index_copy_(0, ids, torch.where(lost, k, d))
        self.steps = steps
        self.theta = (float(V), None if B0 is None else float(B0))
        return float(self.acc.item()) / steps

    # -- backward ---------------------------------------------------------
    def backward(self, grad_out: float = 1.0):
        """(dL/dV, dL/dB0) of the last forward, times grad_out."""
        if not getattr(self, "ckpt", None):
            raise RuntimeError("backward needs a forward with keep=True")
        self._track = False   # the recomputed segments leave the forward's records alone
        V, B0 = self.theta
        steps = self.steps
        ctx, g = self.ctx, self.g
        dev = self.dev
        adj = [AdjointSet(cap, dev) for _, _, cap in self.init]
        lamE = torch.zeros(self.shape + (4,), dtype=torch.float64, device=dev)
        lamphi = torch.zeros(self.shape, dtype=torch.float64, device=dev)
        rhs = torch.zeros(self.shape, dtype=torch.float64, device=dev)
        mu = torch.zeros(self.shape, dtype=torch.float64, device=dev)   # warm start across steps
        lamB = torch.zeros(3, dtype=torch.float64, device=dev)
        lamV = torch.zeros(1, dtype=torch.float64, device=dev)
        plasma = (self.bcnd <= 0)
        B = None if B0 is None or self.B4 is not None else (C.c_double * 3)(0.0, 0.0, B0)
        c = self.case
        scale = grad_out / steps
        starts = sorted(self.ckpt)
        for s0 in reversed(starts):
            s1 = min(s0 + self.K - 1, steps)
            st_c, np_c, phi_c = self.ckpt[s0]
            st = self._snap(st_c)
            self.np_thread.copy_(np_c)
            self.phi.copy_(phi_c)
            states, Es = [], []
            presort = {}    # the state before a sort inside the segment: the previous step's end
            for k in range(s0, s1 + 1):
                if k != s0 and k in self.sorts:
                    presort[k] = self._snap(st)
                    self._apply_sort(st, self.sorts[k])
                states.append(self._snap(st))
                self._fields(V)
                Es.append(self.E4.clone())
                self._push(st, B0)
                if self._collides(k):
                    self._collide(st, k, self.mcc_events[k])
            states.append(self._snap(st))          # after the segment's last step
            for j in range(s1 - s0, -1, -1):
                k = s0 + j
                after, before = presort.get(k + 1, states[j + 1]), states[j]
                # loss term of this step (state after the push and MCC)
                self._flux(after, None, adj[self.flux_species], scale)
                if self._collides(k):
                    # frozen collisions: the velocity a collided particle
                    # leaves with, and a product's state, are constants
                    for s, ev in enumerate(self.mcc_events[k]):
                        a = adj[s].a
                        if ev["idx"].numel():
                            for f in ("lvx", "lvy", "lvz"):
                                a[f].index_fill_(0, ev["idx"], 0.0)
                        if ev["m1"] > ev["m0"]:
                            for v in a.values():
                                v[ev["m0"]:ev["m1"]] = 0.0
                lamE.zero_()
                for s, (m, w) in enumerate(before):
                    p = _pset_of(w, m)
                    a = adj[s].struct()
                    if self.B4 is not None:
                        ctx.check(ctx.lib.lepic_adj_push_map(
                            ctx.h, C.byref(g), C.byref(p), C.byref(a), ptr(Es[j]), ptr(self.B4),
                            int(self.magnetized[s]), self.charge[s], self.mass[s], c.dt,
                            ptr(lamE)))
                    else:
                        ctx.check(ctx.lib.lepic_adj_push(
                            ctx.h, C.byref(g), C.byref(p), C.byref(a), ptr(Es[j]),
                            B if self.magnetized[s] else None, self.charge[s], self.mass[s],
                            c.dt, ptr(lamE), ptr(lamB)))
                ctx.check(ctx.lib.lepic_adj_efield(ctx.h, C.byref(g), ptr(self.bcnd), ptr(lamE),
                                                   ptr(lamphi)))
                # A mu = dL/dphi on plasma nodes, mu = 0 on electrodes, solved
                # by the forward's multigrid at a right-hand side scaled to O(1)
                top = float(torch.where(plasma, lamphi.abs(), 0.0).max().item())
                if top > 0.0:
                    sc = 1.0 / top
                    ctx.check(ctx.lib.lepic_adj_poisson_rhs(ctx.h, C.byref(g), ptr(self.bcnd),
                                                            ptr(lamphi), sc, ptr(rhs)))
                    mu.mul_(sc)          # the previous step's solution as the first guess
                    self._solve(mu, rhs, self.adj_eps, self.adjoint_poisson)
                    mu.mul_(1.0 / sc)
                else:
                    mu.zero_()
                ctx.check(ctx.lib.lepic_adj_bias(ctx.h, C.byref(g), ptr(self.bcnd), self.label,
                                                 ptr(lamphi), ptr(mu), EPS0, ptr(lamV)))
                for s, (m, w) in enumerate(before):
                    p = _pset_of(w, m)
                    a = adj[s].struct()
                    ctx.check(ctx.lib.lepic_adj_deposit(ctx.h, C.byref(g), C.byref(p), C.byref(a),
                                                        ptr(self.kq), ptr(mu), self.charge[s],
                                                        self.Nm[s]))
                if k in self.sorts:
                    # back to the slot order before the sort at the start of
                    # step k; the dropped holes carry no derivative
                    for s, perm in enumerate(self.sorts[k]):
                        adj[s].unpermute_(perm)
            del states, Es, presort
        self.adj = adj
        return float(lamV.item()), (float(lamB[2].item())
                                    if B0 is not None and self.B4 is None else 0.0)


class PicLoss(torch.autograd.Function):
    """L(V, B0) of a DiffPIC run as a torch function: the forward runs the
    steps keeping checkpoints, the backward runs the adjoint kernels.  B0
    None: no uniform field (the case's B map, or none)."""

    @staticmethod
    def forward(ctx, V: torch.Tensor, B0: torch.Tensor | None, sim: DiffPIC, steps: int):
        L = sim.forward(float(V.item()), None if B0 is None else float(B0.item()), steps,
                        keep=True)
        ctx.sim = sim
        ctx.has_b = B0 is not None
        return torch.tensor(L, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gL):
        dV, dB = ctx.sim.backward(float(gL.item()))
        return (torch.tensor(dV, dtype=torch.float64),
                torch.tensor(dB, dtype=torch.float64) if ctx.has_b else None, None, None)
