"""Backup and restart files in the reference's format.

write_backup / read_backup    the particle backup of Src/main.f90:1294-1341
                              (DATA.BAK/particlesN.bak, alternating with
                              DATA.BAK2) and its reader (Src/restart.f90:63-72)
read_external_backups         the merge of backups written with another MPI
                              and OpenMP layout (Src/restart.f90:74-119)
duplicate_or_thin             the np_dup duplication or ran2 thinning
                              (Src/restart.f90:121-167)
restart_simulation            a Simulation from a backup: particles to the
                              device, the initial density deposited as
                              restart_OMP does (Src/restart.f90:194-295)

A backup is a Fortran unformatted sequential file: every WRITE is one
record framed by 4-byte little-endian length markers; a record longer than
2**31 - 9 bytes is split into subrecords whose leading marker is negated
when another subrecord follows and whose trailing marker is negated when
one came before (the convention of gfortran, whose builds of the reference
read and write these files).  Records: the time (one double); np_tot
(ntype x nproc int32, species fastest); then per thread, per species, the
particles as x, y, z, vx, vy, vz doubles particle after particle; after each
thread's species, its negative ions' flag_cex (int32) when the chemistry
has them (tag_neg > 0).  phi.bak (dielectric runs) holds phi over the
ghosted grid, Vgrd.bak (iterative electrode potentials) V(1:ngrid).
"""
from __future__ import annotations

import os
import struct

import numpy as np
import torch

MAX_SUBRECORD = 2 ** 31 - 9


class BackupError(IOError):
    """A backup the reference's reader stops on (err=999)."""


# ---------------------------------------------------------------------------
# unformatted sequential records
# ---------------------------------------------------------------------------

def write_record(f, payload) -> None:
    if isinstance(payload, (bytes, bytearray)):
        data = memoryview(payload)
    else:
        data = memoryview(np.ascontiguousarray(payload).reshape(-1).view(np.uint8))
    n = len(data)
    if n <= MAX_SUBRECORD:
        f.write(struct.pack("<i", n))
        f.write(data)
        f.write(struct.pack("<i", n))
        return
    pos, first = 0, True
    while pos < n:
        m = min(MAX_SUBRECORD, n - pos)
        last = pos + m == n
        f.write(struct.pack("<i", m if last else -m))
        f.write(data[pos:pos + m])
        f.write(struct.pack("<i", m if first else -m))
        pos += m
        first = False


def read_record(f) -> bytes:
    parts = []
    while True:
        head = f.read(4)
        if len(head) < 4:
            raise BackupError("end of file inside the backup")
        (m,) = struct.unpack("<i", head)
        body = f.read(abs(m))
        tail = f.read(4)
        if len(body) < abs(m) or len(tail) < 4:
            raise BackupError("truncated record")
        (t,) = struct.unpack("<i", tail)
        if abs(t) != abs(m):
            raise BackupError("record markers disagree")
        parts.append(body)
        if m >= 0:
            return b"".join(parts)


def _pnum(rank: int) -> str:
    """write(pnum,'(i3)') rank, then pnum(lgh:3): the rank without blanks."""
    if not 0 <= rank <= 999:
        raise ValueError("rank outside 0..999")
    return str(rank)


def backup_path(directory: str, rank: int = 0) -> str:
    return os.path.join(directory, f"particles{_pnum(rank)}.bak")


# ---------------------------------------------------------------------------
# particle backups
# ---------------------------------------------------------------------------

def write_backup(path: str, time: float, species: list, flag_cex: list | None = None) -> None:
    """species[iproc][ptype] = (N, 6) float64 array of x, y, z, vx, vy, vz
    (one thread: species = [[electrons, ions, ...]]); flag_cex[iproc]: int32
    array of the negative-ion species, or None when tag_neg is 0."""
    nproc, ntype = len(species), len(species[0])
    np_tot = np.array([[len(species[i][t]) for i in range(nproc)] for t in range(ntype)],
                      np.int32)                       # (ntype, nproc), species fastest on disk
    with open(path, "wb") as f:
        write_record(f, np.array([time], "<f8"))
        write_record(f, np.asfortranarray(np_tot).ravel(order="F").astype("<i4"))
        for i in range(nproc):
            for t in range(ntype):
                write_record(f, np.ascontiguousarray(species[i][t], "<f8"))
            if flag_cex is not None:
                write_record(f, np.ascontiguousarray(flag_cex[i], "<i4"))


def read_backup(path: str, ntype: int, nproc: int = 1, tag_neg: int = 0):
    """Src/restart.f90:63-72: returns (time, np_tot (ntype, nproc),
    species[iproc][ptype] (N, 6) arrays, flag_cex[iproc] or None)."""
    if not os.path.isfile(path):
        raise BackupError(f"no backup {path}")
    with open(path, "rb") as f:
        rec = read_record(f)
        if len(rec) < 8:
            raise BackupError("time record too short")
        time = struct.unpack("<d", rec[:8])[0]
        rec = read_record(f)
        if len(rec) < 4 * ntype * nproc:
            raise BackupError("np_tot record too short")
        np_tot = np.frombuffer(rec[:4 * ntype * nproc], "<i4").reshape((ntype, nproc), order="F")
        species, cex = [], []
        for i in range(nproc):
            row = []
            for t in range(ntype):
                k = int(np_tot[t, i])
                rec = read_record(f)
                if len(rec) < 48 * k:
                    raise BackupError(f"particle record of species {t + 1} too short")
                row.append(np.frombuffer(rec[:48 * k], "<f8").reshape(k, 6).copy())
            species.append(row)
            if tag_neg > 0:
                k = int(np_tot[tag_neg - 1, i])
                rec = read_record(f)
                cex.append(np.frombuffer(rec[:4 * k], "<i4").copy())
    return time, np.array(np_tot), species, (cex if tag_neg > 0 else None)


def read_external_backups(paths: list, ntype: int, omp_rank_max: int, nproc: int = 1,
                          tag_neg: int = 0):
    """Src/restart.f90:74-119: backups written by other MPI ranks with
    omp_rank_max threads each, dealt round-robin onto nproc threads (appended
    in file, then thread, then species order).  Returns (time, species)."""
    out = [[np.zeros((0, 6)) for _ in range(ntype)] for _ in range(nproc)]
    time = 0.0
    for path in paths:
        time, _, sp, _ = read_backup(path, ntype, omp_rank_max, tag_neg)
        for i in range(omp_rank_max):
            # iproc_tmp counts 1..nproc and wraps (the identity when the
            # thread counts agree)
            tmp = i if omp_rank_max == nproc else i % nproc
            for t in range(ntype):
                out[tmp][t] = np.concatenate([out[tmp][t], sp[i][t]])
    return time, out


def duplicate_or_thin(vxp: np.ndarray, np_dup: float, iseed: list) -> np.ndarray:
    """Src/restart.f90:121-167 for one thread and species: |np_dup| > 1
    repeats the set int(|np_dup|) - 1 more times; |np_dup| < 1 keeps particle
    i when ran2 <= |np_dup| (iseed[0] advanced); otherwise unchanged."""
    a = abs(np_dup)
    if a > 1.0:
        # np_tot becomes int(|np_dup| * np_tot) over int(|np_dup|) copies, so
        # a fractional factor would count particles no copy wrote
        if a != float(int(a)):
            raise BackupError("a duplication factor above 1 must be a whole number")
        k = abs(int(np_dup))
        return np.concatenate([vxp] * k) if len(vxp) else vxp
    if a < 1.0:
        from . import _native
        lib = _native.load()
        import ctypes as C
        seed = C.c_int32(int(iseed[0]))
        draws = np.array([lib.lepic_ran2(C.byref(seed)) for _ in range(len(vxp))])
        iseed[0] = seed.value
        return vxp[draws <= a]
    return vxp


# ---------------------------------------------------------------------------
# Simulation <-> backup
# ---------------------------------------------------------------------------

def save_simulation(sim, directory: str, time: float, rank: int = 0) -> str:
    """One thread's backup of sim's live particles (holes and MCC-killed
    particles left out, as the reference's arrays hold none when it backs up
    after the BC)."""
    os.makedirs(directory, exist_ok=True)
    sp = []
    for p in sim.parts:
        a = p.to_host()
        live = a["dead"] == 0
        sp.append(np.stack([a[k][live] for k in ("x", "y", "z", "vx", "vy", "vz")], 1))
    path = backup_path(directory, rank)
    tag_neg = getattr(sim, "tag_neg", -1)
    cex = [np.zeros(len(sp[tag_neg - 1]), np.int32)] if tag_neg and tag_neg > 0 else None
    write_backup(path, time, [sp], cex)
    return path


def restart_simulation(case, path: str, device=None, np_dup: float = 1.0, rank: int = 0,
                       **sim_kwargs):
    """A Simulation whose particles come from a one-thread backup, with the
    initial density deposited from them (restart_OMP's deposit, which is
    deposit_particle_set_to_np_thread's).  Returns (sim, time)."""
    from . import fields as F
    from . import particles as P
    from .simulation import Simulation
    from .device import node_shape
    dev = torch.device(device if device is not None else "cuda:0")
    ntype = len(case.tracked)
    tag_neg = next((i + 1 for i, s in enumerate(case.tracked) if s.charge < 0 and i > 0), 0)
    time, np_tot, sp, _ = read_backup(path, ntype, 1, tag_neg)
    iseed = [123456 * 1 * (10 * rank + 1)]
    parts = []
    ncells = case.n[0] * case.n[1] * case.n[2]
    base_cap = int(1.3 * abs(case.np_cell) * ncells) + 1024
    for t in range(ntype):
        v = duplicate_or_thin(sp[0][t], np_dup, iseed)
        ps = P.ParticleSet(max(base_cap, len(v) + 1024), dev)
        ps.from_host(*(v[:, c] for c in range(6)))
        parts.append(ps)
    np_thread = torch.zeros((ntype,) + node_shape(case.n), dtype=torch.float64, device=dev)
    kq = F.build_kq(torch.as_tensor(np.ascontiguousarray(case.bcnd, np.int32), device=dev), case.n)
    Nm = case.Nm
    for t, ps in enumerate(parts):
        F.deposit_particle_set_to_np_thread(ps, case.n, case.h, kq, Nm[t], np_thread[t])
    sim = Simulation(case, dev, parts=parts, np_thread=np_thread, **sim_kwargs)
    return sim, time
