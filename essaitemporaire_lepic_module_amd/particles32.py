"""The FP32 particle phase -- the precision form of modules/particles.

ParticleSet32 holds a species as an int32 cell id (-1: no particle), the
position inside the cell in cell widths (fx, fy, fz in [0, 1)) and the
velocity, all FP32: 28 bytes per particle against 53 for the FP64 set.  The
push and gather run in FP32, the deposit in 64-bit fixed point (every
particle's full weight split into 8 integer node parts that sum to it
exactly), so the density is the same for any particle order and the charge
on the grid is exactly conserved.  Each call is one liblepic_hip.so entry
point; the names follow the FP64 module's (move + apply_particle_bc_legacy +
deposit_particle_set_to_np_thread fused as push32, sort_particles_by_cell as
sort32, perform_collisions_step as collisions32).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _native
from .device import context, ptr
from .particles import FIELDS, ParticleSet

FIELDS32 = ("cell", "fx", "fy", "fz", "vx", "vy", "vz")


class ParticleSet32:
    """Device FP32 particle set of one species, two buffers (the sort is out of
    place and swaps them)."""

    def __init__(self, nmax: int = 0, device=None):
        self.device = torch.device(device if device is not None else "cuda:0")
        self.n = 0
        self.nmax = 0
        self._buf = [None, None]
        self._cur = 0
        if nmax:
            self.allocate(nmax)

    def _alloc(self, cap: int) -> dict:
        # padded to whole 16-byte vectors so every array stays 16-byte aligned
        cap4 = (cap + 3) // 4 * 4
        d = {"cell": torch.full((cap4,), -1, dtype=torch.int32, device=self.device)}
        for f in FIELDS32[1:]:
            d[f] = torch.zeros(cap4, dtype=torch.float32, device=self.device)
        return d

    def allocate(self, nmax: int):
        nmax = max(0, int(nmax))
        self._buf = [self._alloc(nmax), self._alloc(nmax)]
        self._cur = 0
        self.n = 0
        self.nmax = nmax

    def ensure_capacity(self, needed: int):
        """Grow by doubling, keeping the particles in use."""
        needed = int(needed)
        if needed <= self.nmax:
            return
        new = max(needed, 2 * self.nmax, 16)
        cur = self._alloc(new)
        old = self._buf[self._cur]
        if old is not None and self.n:
            for k in cur:
                cur[k][: self.n].copy_(old[k][: self.n])
        self._buf = [cur, self._alloc(new)]
        self._cur = 0
        self.nmax = new

    def swap(self):
        self._cur ^= 1

    def __getattr__(self, name):
        if name in FIELDS32:
            return self._buf[self._cur][name][: self.n]
        raise AttributeError(name)

    def pset(self, other: bool = False) -> _native.PSet32:
        b = self._buf[self._cur ^ (1 if other else 0)]
        p = _native.PSet32()
        for f in FIELDS32:
            setattr(p, f, b[f].data_ptr())
        p.n = self.n
        p.cap = self.nmax
        return p

    # -- host transfer ---------------------------------------------------
    def from_host(self, cell, fx, fy, fz, vx, vy, vz):
        m = len(cell)
        if m > self.nmax:
            self.allocate(m)
        b = self._buf[self._cur]
        b["cell"].fill_(-1)
        b["cell"][:m].copy_(torch.as_tensor(np.asarray(cell, np.int32)))
        for f, v in zip(FIELDS32[1:], (fx, fy, fz, vx, vy, vz)):
            b[f][:m].copy_(torch.as_tensor(np.asarray(v, np.float32)))
        self.n = m
        return self

    def to_host(self) -> dict:
        b = self._buf[self._cur]
        return {k: b[k][: self.n].cpu().numpy() for k in FIELDS32}

    # -- layout conversions ---------------------------------------------
    @classmethod
    def from_pset64(cls, part: ParticleSet, n, h, cap: int | None = None) -> tuple:
        """The FP64 set converted slot for slot (lepic_pset_to32); returns the
        set and the count of particles outside [0, L) (made holes)."""
        q = cls(cap or max(part.nmax, part.n, 16), part.device)
        ctx = context(part.device)
        src, dst = part.pset(), q.pset()
        out = C.c_int64(0)
        ctx.check(ctx.lib.lepic_pset_to32(ctx.h, C.byref(_native.grid(n, h)), C.byref(src),
                                          C.byref(dst), C.byref(out)))
        q.n = int(dst.n)
        return q, int(out.value)

    def to_pset64(self, n, h) -> ParticleSet:
        """The FP64 set of the same particles, slot for slot (holes dead = 2)."""
        p = ParticleSet(max(self.nmax, 16), self.device)
        ctx = context(self.device)
        src, dst = self.pset(), p.pset()
        ctx.check(ctx.lib.lepic_pset_from32(ctx.h, C.byref(_native.grid(n, h)), C.byref(src),
                                            C.byref(dst)))
        p.n = int(dst.n)
        return p


def bfield32(mode: int = 0, B=(0.0, 0.0, 0.0), B4f: torch.Tensor | None = None,
             magnetized: bool = True) -> _native.BField32:
    """Magnetic field of the FP32 push: 0 none, 1 uniform B, 2 a map on the
    field grid (float32, 4 per node: Bx, By, Bz, |B|)."""
    bf = _native.BField32()
    bf.mode = int(mode)
    for i in range(3):
        bf.B[i] = float(B[i])
    bf.B4 = ptr(B4f) if B4f is not None else None
    bf.magnetized = 1 if magnetized else 0
    bf._keep = B4f
    return bf


def round_field4(F4: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
    """A field of 4 doubles per node rounded to float32 (E for the FP32
    gather, B maps)."""
    if out is None:
        out = torch.empty(F4.shape, dtype=torch.float32, device=F4.device)
    ctx = context(F4.device)
    ctx.check(ctx.lib.lepic_round_field4(ctx.h, ptr(F4), F4.numel() // 4, ptr(out)))
    return out


def push32(part: ParticleSet32, n, h, E4f, bf: _native.BField32, q, m, dt, bcnd,
           bp: _native.BCParams, frac_bits: int, dep: torch.Tensor | None, p_mac, P_loss,
           lost_acc: torch.Tensor, escaped: torch.Tensor):
    """Gather + push + wall BC + fixed-point deposit in FP32, in place (lost
    particles become holes, counted in lost_acc; survivors outside the cells
    in escaped).  dep: int64 node sums of this species (None: no deposit)."""
    ctx = context(part.device)
    p = part.pset()
    ctx.check(ctx.lib.lepic_push32(ctx.h, C.byref(_native.grid(n, h)), C.byref(p), ptr(E4f),
                                   C.byref(bf), float(q), float(m), float(dt), ptr(bcnd),
                                   C.byref(bp), int(frac_bits), ptr(dep), ptr(p_mac),
                                   ptr(P_loss), ptr(lost_acc), ptr(escaped)))


def deposit32(part: ParticleSet32, n, h, Nm: float, frac_bits: int, dep: torch.Tensor):
    """Fixed-point deposit of the set, added to dep (int64 per node)."""
    ctx = context(part.device)
    p = part.pset()
    ctx.check(ctx.lib.lepic_deposit32(ctx.h, C.byref(_native.grid(n, h)), C.byref(p), float(Nm),
                                      int(frac_bits), ptr(dep)))


def fixed_density(n, h, dep: torch.Tensor, kq: torch.Tensor, frac_bits: int,
                  np_local: torch.Tensor):
    """np_local += kq * dep * 2^-frac_bits."""
    ctx = context(dep.device)
    ctx.check(ctx.lib.lepic_fixed_density(ctx.h, C.byref(_native.grid(n, h)), ptr(dep), ptr(kq),
                                          int(frac_bits), ptr(np_local)))


def sort32(part: ParticleSet32, n, h):
    """Stable sort by cell dropping holes (sort_particles_by_cell,
    mod_particle_sorting.f90:88-181)."""
    ctx = context(part.device)
    src, dst = part.pset(), part.pset(other=True)
    ctx.check(ctx.lib.lepic_sort32(ctx.h, C.byref(_native.grid(n, h)), C.byref(src),
                                   C.byref(dst)))
    part.swap()
    part.n = int(dst.n)


def collisions32(parts: list, n, h, np_red: torch.Tensor, tables, ns_coll: int, dt: float,
                 Pcoll: torch.Tensor, philox_key: int, philox_counter: int,
                 dep: torch.Tensor | None = None, frac_bits=None) -> dict:
    """perform_collisions_step (mod_collisions.f90:214-739) on FP32 sets,
    PHILOX mode.  dep ([ntype][nodes] int64) receives the fixed-point deposit
    of appended products and loses that of killed particles."""
    ctx = context(parts[0].device)
    nt = len(parts)
    arr = (_native.PSet32 * nt)(*[p.pset() for p in parts])
    st = _native.MCCStats()
    fb = (C.c_int32 * max(nt, 1))(*[int(b) for b in (frac_bits or [0] * nt)])
    ctx.check(ctx.lib.lepic_collisions32(ctx.h, C.byref(_native.grid(n, h)), arr, ptr(np_red),
                                         C.byref(tables.c), int(ns_coll), float(dt),
                                         C.c_uint64(philox_key), C.c_uint64(philox_counter),
                                         ptr(Pcoll), C.byref(st), ptr(dep), fb))
    for p, s in zip(parts, arr):
        p.n = int(s.n)
    return {"Nc": [st.Nc[i] for i in range(nt)], "accepted": [st.accepted[i] for i in range(nt)],
            "appended": [st.appended[i] for i in range(nt)],
            "skipped_dead": [st.skipped_dead[i] for i in range(nt)]}


__all__ = ["ParticleSet32", "bfield32", "round_field4", "push32", "deposit32", "fixed_density",
           "sort32", "collisions32", "FIELDS", "FIELDS32"]
