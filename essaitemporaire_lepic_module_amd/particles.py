"""Particle side of the step -- the call surface of modules/particles.

ParticleSet (mod_particles.f90:7-34) lives in HBM as structure-of-arrays:
x, y, z, vx, vy, vz (FP64), cell (int32) and dead (int8, the reference's
flag_dead).  Each set owns two buffers; out-of-place passes (sort, wall BC)
write the other buffer and swap, so the caller sees the reference's in-place
semantics.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _native
from .device import as_device, context, dbl_array, ptr

FIELDS = ("x", "y", "z", "vx", "vy", "vz")


class ParticleSet:
    """Device particle set of one species (mod_particles.f90:7-34)."""

    def __init__(self, nmax: int = 0, device=None):
        self.device = torch.device(device if device is not None else "cuda:0")
        self.n = 0
        self.nmax = 0
        self._buf = [None, None]
        self._cur = 0
        self.cell_count = None
        self.cell_start = None
        if nmax:
            self.allocate_pset(nmax)

    # -- storage ---------------------------------------------------------
    def _alloc(self, cap: int) -> dict:
        d = {f: torch.zeros(cap, dtype=torch.float64, device=self.device) for f in FIELDS}
        d["cell"] = torch.zeros(cap, dtype=torch.int32, device=self.device)
        d["dead"] = torch.zeros(cap, dtype=torch.int8, device=self.device)
        return d

    def allocate_pset(self, nmax: int):
        """allocate_pset (mod_particles.f90:39-72).  The second buffer (the
        target of out-of-place passes) is allocated at its first use."""
        nmax = max(0, int(nmax))
        self._buf = [self._alloc(nmax), None]
        self._cur = 0
        self.n = 0
        self.nmax = nmax

    def ensure_capacity(self, needed: int):
        """Grow by doubling, keeping the live particles (mod_particles.f90:74-151)."""
        needed = int(needed)
        if needed <= self.nmax:
            return
        new = max(needed, 2 * self.nmax, 16)
        cur = self._alloc(new)
        old = self._buf[self._cur]
        if old is not None and self.n:
            for k in cur:
                cur[k][: self.n].copy_(old[k][: self.n])
        self._buf = [cur, None]
        self._cur = 0
        self.nmax = new

    def _buffer(self, i: int) -> dict:
        if self._buf[i] is None:
            self._buf[i] = self._alloc(self.nmax)
        return self._buf[i]

    def release(self):
        """Drops both buffers (the set is empty afterwards)."""
        self._buf = [None, None]
        self._cur = 0
        self.n = 0
        self.nmax = 0

    def swap(self):
        self._buffer(self._cur ^ 1)
        self._cur ^= 1

    def arrays(self, which: int | None = None) -> dict:
        return self._buffer(self._cur if which is None else which)

    def __getattr__(self, name):
        if name in FIELDS or name in ("cell", "dead"):
            return self._buffer(self._cur)[name][: self.n]
        raise AttributeError(name)

    def pset(self, other: bool = False) -> _native.PSet:
        b = self._buffer(self._cur ^ (1 if other else 0))
        p = _native.PSet()
        for f in FIELDS + ("cell", "dead"):
            setattr(p, f, b[f].data_ptr())
        p.n = self.n
        p.cap = self.nmax
        return p

    # -- host transfer ---------------------------------------------------
    def from_host(self, x, y, z, vx, vy, vz, dead=None, cell=None):
        m = len(x)
        if m > self.nmax:
            self.allocate_pset(m)
        b = self._buffer(self._cur)
        for f, v in zip(FIELDS, (x, y, z, vx, vy, vz)):
            b[f][:m].copy_(torch.as_tensor(np.asarray(v, np.float64)))
        b["dead"].zero_()
        if dead is not None:
            b["dead"][:m].copy_(torch.as_tensor(np.asarray(dead, np.int8)))
        if cell is not None:
            b["cell"][:m].copy_(torch.as_tensor(np.asarray(cell, np.int32)))
        self.n = m
        return self

    def to_host(self) -> dict:
        b = self._buffer(self._cur)
        return {k: b[k][: self.n].cpu().numpy() for k in FIELDS + ("cell", "dead")}


# ---------------------------------------------------------------------------
# a2: sort (mod_particle_sorting.f90)
# ---------------------------------------------------------------------------

def cell_index_from_position(x, y, z, h, n) -> int:
    """1-based cell id of one position (mod_particle_sorting.f90:16-48)."""
    ix = max(1, min(int(n[0]), int(x / h[0]) + 1))
    iy = max(1, min(int(n[1]), int(y / h[1]) + 1))
    iz = max(1, min(int(n[2]), int(z / h[2]) + 1))
    return (ix - 1) + int(n[0]) * ((iy - 1) + int(n[1]) * (iz - 1)) + 1


def sort_particles_by_cell(part: ParticleSet, n, h, cell_tables: bool = True):
    """Stable counting sort by cell (mod_particle_sorting.f90:88-181).

    On return the arrays are sorted, part.cell holds 0-based cell ids and,
    with cell_tables, part.cell_count / part.cell_start (0-based) are valid."""
    ctx = context(part.device)
    g = _native.grid(n, h)
    ncells = int(n[0]) * int(n[1]) * int(n[2])
    cc = cs = None
    if cell_tables:
        if part.cell_count is None or part.cell_count.numel() != ncells:
            part.cell_count = torch.zeros(ncells, dtype=torch.int32, device=part.device)
            part.cell_start = torch.zeros(ncells, dtype=torch.int32, device=part.device)
        cc, cs = part.cell_count, part.cell_start
    src, dst = part.pset(), part.pset(other=True)
    ctx.check(ctx.lib.lepic_sort_particles_by_cell(ctx.h, C.byref(g), C.byref(src), C.byref(dst),
                                                   ptr(cc), ptr(cs)))
    part.swap()


# ---------------------------------------------------------------------------
# a3/a4: movers
# ---------------------------------------------------------------------------

def move_particles_electrostatic(part: ParticleSet, n, h, E4: torch.Tensor, q: float, m: float,
                                 dt: float):
    """Gather + leapfrog push in place (mod_particleMover.f90:72-156).
    E4: (n3+3, n2+3, n1+3, 4) FP64, components (Ex, Ey, Ez, pad)."""
    ctx = context(part.device)
    g = _native.grid(n, h)
    p = part.pset()
    ctx.check(ctx.lib.lepic_move_electrostatic(ctx.h, C.byref(g), C.byref(p), ptr(E4), q, m, dt))


def bfield(mode: int = 0, B=(0.0, 0.0, 0.0), B4: torch.Tensor | None = None,
           magnetized: bool = True, nB=None, hB=None) -> _native.BField:
    """Magnetic field of the push: mode 0 none, 1 uniform B, 2 map B4 on the
    field grid, 3 map B4 on its own grid of nB cells of size hB (B4 then has
    (nB3+3, nB2+3, nB1+3, 4) doubles)."""
    bf = _native.BField()
    bf.mode = int(mode)
    for i in range(3):
        bf.B[i] = float(B[i])
    bf.B4 = ptr(B4) if B4 is not None else None
    bf.magnetized = 1 if magnetized else 0
    if mode == 3:
        if nB is None or hB is None:
            raise ValueError("a map on its own grid needs nB and hB")
        if B4 is not None and tuple(B4.shape) != (nB[2] + 3, nB[1] + 3, nB[0] + 3, 4):
            raise ValueError("B4 must be (nB3+3, nB2+3, nB1+3, 4)")
        for i in range(3):
            bf.nB[i] = int(nB[i])
            bf.hB[i] = float(hB[i])
    bf._keep = B4
    return bf


def move_particles_boris(part: ParticleSet, n, h, E4, bf: _native.BField, q: float, m: float,
                         dt: float):
    """Boris (or leapfrog where B is off) push in place (Src/part_expmover.f90:51-246)."""
    ctx = context(part.device)
    g = _native.grid(n, h)
    p = part.pset()
    ctx.check(ctx.lib.lepic_move_boris(ctx.h, C.byref(g), C.byref(p), ptr(E4), C.byref(bf), q, m, dt))


# ---------------------------------------------------------------------------
# a5: wall BC
# ---------------------------------------------------------------------------

def bc_params(n, h, flag_pbc, flag_nmn, ptype, tag_neg, qmacro, mass, Nm, ngrid,
              flag_die: int = 0, dtype: torch.Tensor | None = None,
              sum_q_xz: torch.Tensor | None = None,
              sum_q_yz: torch.Tensor | None = None) -> _native.BCParams:
    """Wall-BC parameters; with flag_die, dtype (device int32 over labels
    0..ngrid) and this species' surface-charge planes sum_q_xz (n3+3, n1+3)
    and sum_q_yz (n3+3, n2+3, 2)."""
    bp = _native.BCParams()
    bp.xmax, bp.ymax, bp.zmax = (float(n[d]) * float(h[d]) for d in range(3))
    bp.flag_pbc, bp.flag_nmn = int(flag_pbc), int(flag_nmn)
    bp.ptype, bp.tag_neg = int(ptype), int(tag_neg)
    bp.qmacro, bp.mass, bp.Nm = float(qmacro), float(mass), float(Nm)
    bp.ngrid = int(ngrid)
    bp.flag_die = int(flag_die)
    if flag_die:
        if dtype is None or sum_q_xz is None or sum_q_yz is None:
            raise ValueError("flag_die needs dtype, sum_q_xz and sum_q_yz")
        if dtype.dtype != torch.int32 or dtype.numel() != int(ngrid) + 1:
            raise ValueError("dtype must be int32 over labels 0..ngrid")
        n1, n2, n3 = (int(v) for v in n)
        if (tuple(sum_q_xz.shape) != (n3 + 3, n1 + 3) or tuple(sum_q_yz.shape) != (n3 + 3, n2 + 3, 2)
                or sum_q_xz.dtype != torch.float64 or sum_q_yz.dtype != torch.float64
                or not sum_q_xz.is_contiguous() or not sum_q_yz.is_contiguous()):
            raise ValueError("sum_q_xz / sum_q_yz: contiguous float64 (n3+3, n1+3), (n3+3, n2+3, 2)")
        bp.dtype, bp.sum_q_xz, bp.sum_q_yz = ptr(dtype), ptr(sum_q_xz), ptr(sum_q_yz)
        bp._keep = (dtype, sum_q_xz, sum_q_yz)
    return bp


def apply_particle_bc_legacy(part: ParticleSet, n, h, bcnd: torch.Tensor, flag_pbc: int,
                             flag_nmn: int, ptype: int, tag_neg: int, flag_die: int, qmacro: float,
                             p_mac_boundary: torch.Tensor, mass_species: float,
                             P_loss_wall: torch.Tensor, Nm_species: float,
                             dtype: torch.Tensor | None = None,
                             sum_q_xz: torch.Tensor | None = None,
                             sum_q_yz: torch.Tensor | None = None) -> int:
    """Wall loss, dielectric surface charging, Neumann reflection, periodic
    wrap and stable compaction (mod_particleBC.f90:42-259).  p_mac_boundary:
    (ngrid+1, 2) FP64 device tensor indexed by electrode label; P_loss_wall:
    1-element FP64 tensor; with flag_die, the domain's dtype (labels
    0..ngrid) and the species' surface-charge planes.  Returns the number of
    particles removed."""
    ctx = context(part.device)
    g = _native.grid(n, h)
    ngrid = p_mac_boundary.shape[0] - 1
    bp = bc_params(n, h, flag_pbc, flag_nmn, ptype, tag_neg, qmacro, mass_species, Nm_species, ngrid,
                   flag_die, dtype, sum_q_xz, sum_q_yz)
    src, dst = part.pset(), part.pset(other=True)
    lost = C.c_int64(0)
    ctx.check(ctx.lib.lepic_apply_bc(ctx.h, C.byref(g), C.byref(src), C.byref(dst), ptr(bcnd),
                                     C.byref(bp), ptr(p_mac_boundary), ptr(P_loss_wall),
                                     C.byref(lost)))
    part.swap()
    part.n = int(dst.n)
    return int(lost.value)


# ---------------------------------------------------------------------------
# fused particle phase
# ---------------------------------------------------------------------------

def push_bc_deposit(part: ParticleSet, n, h, E4, bf: _native.BField, q, m, dt, bcnd,
                    bp: _native.BCParams, kq, np_local, p_mac, P_loss) -> int:
    """Mover + apply_particle_bc_legacy + deposit in one pass (np_local may be
    None to skip the deposit).  Returns the number of particles removed."""
    ctx = context(part.device)
    g = _native.grid(n, h)
    src, dst = part.pset(), part.pset(other=True)
    lost = C.c_int64(0)
    ctx.check(ctx.lib.lepic_push_bc_deposit(ctx.h, C.byref(g), C.byref(src), C.byref(dst), ptr(E4),
                                            C.byref(bf), q, m, dt, ptr(bcnd), C.byref(bp), ptr(kq),
                                            ptr(np_local), ptr(p_mac), ptr(P_loss), C.byref(lost)))
    part.swap()
    part.n = int(dst.n)
    return int(lost.value)


def push_bc_deposit_inplace(part: ParticleSet, n, h, E4, bf: _native.BField, q, m, dt, bcnd,
                            bp: _native.BCParams, kq, np_local, p_mac, P_loss,
                            lost_acc: torch.Tensor | None = None, sync: bool = True):
    """Production particle phase in place: lost particles become holes
    (dead = 2), removed later by sort_compact.  np_local None skips the
    deposit.  lost_acc (1-element int64 device tensor) accumulates the
    losses on the device; with sync the call waits for the stream and
    returns this call's losses, otherwise it returns None."""
    ctx = context(part.device)
    g = _native.grid(n, h)
    p = part.pset()
    lost = C.c_int64(0)
    ctx.check(ctx.lib.lepic_push_bc_deposit_inplace(ctx.h, C.byref(g), C.byref(p), ptr(E4),
                                                    C.byref(bf), q, m, dt, ptr(bcnd), C.byref(bp),
                                                    ptr(kq), ptr(np_local), ptr(p_mac), ptr(P_loss),
                                                    ptr(lost_acc), C.byref(lost) if sync else None))
    return int(lost.value) if sync else None


def bin_compact(part: ParticleSet, n, h):
    """Group particles by cell (cell order of sort_particles_by_cell, order
    inside a cell unspecified) and drop holes and MCC-killed particles."""
    ctx = context(part.device)
    g = _native.grid(n, h)
    src, dst = part.pset(), part.pset(other=True)
    ctx.check(ctx.lib.lepic_bin_compact(ctx.h, C.byref(g), C.byref(src), C.byref(dst)))
    part.swap()
    part.n = int(dst.n)


def sort_compact(part: ParticleSet, n, h):
    """Stable cell sort that drops holes and MCC-killed particles."""
    ctx = context(part.device)
    g = _native.grid(n, h)
    src, dst = part.pset(), part.pset(other=True)
    ctx.check(ctx.lib.lepic_sort_compact(ctx.h, C.byref(g), C.byref(src), C.byref(dst)))
    part.swap()
    part.n = int(dst.n)


# ---------------------------------------------------------------------------
# a10: MCC
# ---------------------------------------------------------------------------

class ReactionTables:
    """Device copy of the reaction tables (layout of lepic_chem)."""

    def __init__(self, t: dict, Nm, device=None):
        dev = torch.device(device if device is not None else "cuda:0")
        self.t = t
        self.sig_list = as_device(np.asarray(t["sig_list"], np.int32), torch.int32, dev)
        self.sig_Er = as_device(np.asarray(t["sig_Er"], np.float64), torch.float64, dev)
        self.sig = as_device(np.asarray(t["sig"], np.float64), torch.float64, dev)
        self.sig_Eth = as_device(np.asarray(t["sig_Eth"], np.float64), torch.float64, dev)
        self.col_info = as_device(np.asarray(t["col_info"], np.int32), torch.int32, dev)
        self.sigv_mx = as_device(np.asarray(t["sigv_mx"], np.float64), torch.float64, dev)
        c = _native.Chem()
        c.ntype_tracked = int(t["ntype_tracked"])
        c.ntype_total = int(t["ntype_total"])
        c.mass = dbl_array(t["mass"])
        c.vt0 = dbl_array(t["vt0"])
        c.Nm = dbl_array(Nm)
        c.neutral_density = dbl_array(t["neutral_density"])
        c.nu_uplim = dbl_array(t["nu_uplim"])
        for i, v in enumerate(t["p_ncol"]):
            c.p_ncol[i] = int(v)
        c.ncol, c.npt = int(t["ncol"]), int(t["npt"])
        c.sig_list, c.sig_Er, c.sig = ptr(self.sig_list), ptr(self.sig_Er), ptr(self.sig)
        c.sig_Eth, c.col_info, c.sigv_mx = ptr(self.sig_Eth), ptr(self.col_info), ptr(self.sigv_mx)
        self.c = c


def perform_collisions_step(parts: list, n, h, np_red: torch.Tensor, tables: ReactionTables,
                            ns_coll: int, dt: float, iseed: list | None, Pcoll: torch.Tensor,
                            mode: int = _native.MCC_REPLAY_RAN2, philox_key: int = 0,
                            philox_counter: int = 0, np_dep: torch.Tensor | None = None,
                            kq: torch.Tensor | None = None) -> dict:
    """Null-collision MCC (mod_collisions.f90:214-739) on the tracked species.

    mode MCC_REPLAY_RAN2 replays the reference's sequential chain with the
    ran2 seed iseed[0] (updated in place); MCC_PHILOX is the parallel
    production path.  np_dep (with kq): the step's density, deposited before
    MCC; products are added to it and killed particles taken out, so that it
    equals the deposit after MCC.  Returns per-species event counts."""
    ctx = context(parts[0].device)
    g = _native.grid(n, h)
    nt = len(parts)
    arr = (_native.PSet * nt)(*[p.pset() for p in parts])
    seed = C.c_int32(int(iseed[0]) if iseed else 0)
    st = _native.MCCStats()
    args = (ctx.h, C.byref(g), arr, ptr(np_red), C.byref(tables.c), int(ns_coll), float(dt),
            int(mode), C.byref(seed), C.c_uint64(philox_key), C.c_uint64(philox_counter),
            ptr(Pcoll), C.byref(st))
    if np_dep is not None:
        if kq is None:
            raise ValueError("np_dep needs kq")
        ctx.check(ctx.lib.lepic_collisions_deposited(*args, ptr(np_dep), ptr(kq)))
    else:
        ctx.check(ctx.lib.lepic_collisions(*args))
    for p, s in zip(parts, arr):
        p.n = int(s.n)
    if iseed is not None and mode == _native.MCC_REPLAY_RAN2:
        iseed[0] = seed.value
    return {"Nc": [st.Nc[i] for i in range(nt)], "accepted": [st.accepted[i] for i in range(nt)],
            "err_coll": [st.err_coll[i] for i in range(nt)],
            "skipped_dead": [st.skipped_dead[i] for i in range(nt)],
            "appended": [st.appended[i] for i in range(nt)]}


# ---------------------------------------------------------------------------
# f4: electron heating
# ---------------------------------------------------------------------------

def heating_params(nudt: float, Pabs: float, nu_h: float, xl_pow: float, xr_pow: float,
                   ymax: float, zmax: float, Nm: float, mass: float, vt0: float,
                   flag_circxh: int = 0, flag_ahp: int = 0, R_ahp_moments: float = 0.0,
                   R_ahp: float = 0.0) -> _native.Heating:
    hp = _native.Heating()
    hp.nudt, hp.Pabs, hp.nu_h = float(nudt), float(Pabs), float(nu_h)
    hp.xl_pow, hp.xr_pow = float(xl_pow), float(xr_pow)
    hp.flag_circxh, hp.flag_ahp = int(flag_circxh), int(flag_ahp)
    hp.R_ahp_moments, hp.R_ahp = float(R_ahp_moments), float(R_ahp)
    hp.ymax, hp.zmax, hp.Nm, hp.mass, hp.vt0 = (float(v) for v in (ymax, zmax, Nm, mass, vt0))
    return hp


def apply_electron_heating(part: ParticleSet, n, h, hp: _native.Heating, iseed: list | None,
                           P_heat: torch.Tensor, mode: int = _native.MCC_REPLAY_RAN2,
                           philox_key: int = 0, philox_counter: int = 0,
                           stats: bool = True) -> dict | None:
    """compute_heating_region_moments, update_heating_vt and
    apply_electron_heating (mod_state.f90:341-373, 524-612; mod_heating.f90:12-82)
    on the electrons.  MCC_REPLAY_RAN2 consumes the reference's ran2 stream
    iseed[0] (updated in place); MCC_PHILOX draws per particle.  P_heat
    (device, 1 double) accumulates the absorbed energy.  Returns the window
    count, its kinetic energy and the vt used (stats=False: None, and in
    MCC_PHILOX mode the call then does not wait for the device)."""
    ctx = context(part.device)
    seed = C.c_int32(int(iseed[0]) if iseed else 0)
    out = (C.c_double * 3)() if stats else None
    ps = part.pset()
    ctx.check(ctx.lib.lepic_electron_heating(ctx.h, C.byref(_native.grid(n, h)), C.byref(ps),
                                             C.byref(hp), int(mode), C.byref(seed),
                                             C.c_uint64(philox_key), C.c_uint64(philox_counter),
                                             ptr(P_heat), out))
    if iseed is not None and mode == _native.MCC_REPLAY_RAN2:
        iseed[0] = seed.value
    return {"Nh": int(out[0]), "sum_Ek": out[1], "vt": out[2]} if stats else None


# ---------------------------------------------------------------------------
# a11/a13: RNG and loaders
# ---------------------------------------------------------------------------

def ran2(iseed: list) -> float:
    """One ran2 draw; iseed[0] is updated (mod_RNG.f90:28-45)."""
    lib = _native.load()
    s = C.c_int32(int(iseed[0]))
    v = lib.lepic_ran2(C.byref(s))
    iseed[0] = s.value
    return v


def seed_initialization(iproc: int, mpi_rank: int = 0) -> int:
    return int(_native.load().lepic_seed(int(iproc), int(mpi_rank)))


def load_particles_modular(n, h, bcnd_host: np.ndarray, kq_host: np.ndarray, np_cell: int,
                           vt0, Nm, ni0, iseed: list, x_load: float = 0.0, cap: int | None = None,
                           device=None):
    """Reference loader (mod_particle_loader.f90:14-246, one virtual thread) run
    on the host with ran2; returns device ParticleSets and the initial
    np_thread (ntype, nodes...) as a device tensor."""
    lib = _native.load()
    nt = len(vt0)
    bc = np.ascontiguousarray(bcnd_host, np.int32)
    kq = np.ascontiguousarray(kq_host, np.float64)
    ncells = int(n[0]) * int(n[1]) * int(n[2])
    cap = int(cap) if cap else int(1.3 * np_cell * ncells) + 32
    host = []
    psets = (_native.PSet * nt)()
    for t in range(nt):
        a = {f: np.zeros(cap) for f in FIELDS}
        a["cell"] = np.zeros(cap, np.int32)
        a["dead"] = np.zeros(cap, np.int8)
        host.append(a)
        for f in FIELDS + ("cell", "dead"):
            setattr(psets[t], f, a[f].ctypes.data)
        psets[t].n, psets[t].cap = 0, cap
    np_thread = np.zeros((nt,) + bc.shape)
    seed = (C.c_int32 * 1)(int(iseed[0]))
    g = _native.grid(n, h)
    st = lib.lepic_load_particles_host(
        C.byref(g), bc.ctypes.data, kq.ctypes.data, float(x_load), int(np_cell), nt,
        dbl_array(vt0, nt), dbl_array(Nm, nt), dbl_array(ni0, nt), seed, psets, np_thread.ctypes.data)
    if st != 0:
        raise _native.LepicError(st, "load_particles_modular failed")
    iseed[0] = seed[0]
    dev = torch.device(device if device is not None else "cuda:0")
    parts = []
    for t in range(nt):
        m = int(psets[t].n)
        ps = ParticleSet(max(cap, 16), dev)
        a = host[t]
        ps.from_host(a["x"][:m], a["y"][:m], a["z"][:m], a["vx"][:m], a["vy"][:m], a["vz"][:m])
        parts.append(ps)
    return parts, torch.as_tensor(np_thread, device=dev)


def load_uniform_device(n, h, bcnd: torch.Tensor, np_cell: int, vt0, key: int, cap: int,
                        device=None):
    """Device loader: np_cell particles per loadable cell, Maxwellian velocities
    from Philox (statistically the reference loader)."""
    dev = torch.device(device if device is not None else "cuda:0")
    ctx = context(dev)
    nt = len(vt0)
    parts = [ParticleSet(cap, dev) for _ in range(nt)]
    arr = (_native.PSet * nt)(*[p.pset() for p in parts])
    g = _native.grid(n, h)
    ctx.check(ctx.lib.lepic_load_uniform_device(ctx.h, C.byref(g), ptr(bcnd), int(np_cell), nt,
                                                dbl_array(vt0, nt), C.c_uint64(key), arr))
    for p, s in zip(parts, arr):
        p.n = int(s.n)
    return parts
