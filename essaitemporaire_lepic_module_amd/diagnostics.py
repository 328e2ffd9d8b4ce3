"""Diagnostics of the step: plane moments of the particles, time-averaged
density and potential planes, the power and wall-current summary, and the
.mco plane files.

compute_particle_plane_moments_species
        modules/particles/mod_planeMoments.f90:13-308, on the device
        (lepic_plane_moments)
Diagnostics.accumulate_2d_averages / reset_2d_averages
        modules/core/mod_state.f90:636-667, modules/simulation/mod_simulation.f90:224-240
Diagnostics.power_diagnostics / print_diagnostics
        modules/simulation/mod_simulation.f90:574-694
Diagnostics.output_step, write_initial_diagnostics
        modules/simulation/mod_simulation.f90:99-146, 243-319
write_plane_* / write_*_planes
        modules/io/mod_output_2d.f90:13-258

The cadences are the modular driver's: moments and plane sums on steps with
mod(istep, nb_step_averaging) == 1, after the field solve and before the push
(mod_simulation.f90:408-416); files and the reset of the sums on steps with
mod(istep, nsav) == 1, after the deposit (:455-460); the printed summary
before such a step (:486-493).  Plane arrays are torch tensors on the device
with the plane's second axis first: an xy plane is (n2+3, n1+3).
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np
import torch

from . import _native
from .config import QE
from .device import context, ptr

NP_AVG, TP_AVG, U1_AVG, U2_AVG, U3_AVG = range(5)   # mod_planeMoments.f90:31


# ---------------------------------------------------------------------------
# plane moments on the device
# ---------------------------------------------------------------------------

def plane_shapes(n):
    n1, n2, n3 = (int(v) for v in n)
    return (n2 + 3, n1 + 3), (n3 + 3, n1 + 3), (n3 + 3, n2 + 3)


def compute_particle_plane_moments_species(part, n, h, dt: float, mass: float, ix_plane: int,
                                           iy_plane: int, iz_plane: int, data_xy: torch.Tensor,
                                           data_xz: torch.Tensor, data_yz: torch.Tensor):
    """Plane moments of one species (mod_planeMoments.f90:13-308) into
    data_xy (n2+3, n1+3, 5), data_xz (n3+3, n1+3, 5), data_yz (n3+3, n2+3, 5):
    np and u1..u3 overwritten where a node has weight, T accumulated."""
    sxy, sxz, syz = plane_shapes(n)
    for d, s in ((data_xy, sxy), (data_xz, sxz), (data_yz, syz)):
        if tuple(d.shape) != s + (5,) or d.dtype != torch.float64 or not d.is_contiguous():
            raise ValueError(f"plane data must be contiguous float64 of shape {s + (5,)}")
    ctx = context(part.device)
    ctx.check(ctx.lib.lepic_plane_moments(ctx.h, C.byref(_native.grid(n, h)), C.byref(part.pset()),
                                          float(dt), float(mass), int(ix_plane), int(iy_plane),
                                          int(iz_plane), ptr(data_xy), ptr(data_xz), ptr(data_yz)))


# ---------------------------------------------------------------------------
# Fortran edit descriptors of the writers
# ---------------------------------------------------------------------------

def fortran_es(v: float, w: int, d: int) -> str:
    """ESw.d: one digit before the point, d after, exponent E+dd (or +ddd
    without the E when it needs three digits), right-justified in w."""
    v = float(v)
    if math.isnan(v):
        s = "NaN"
    elif math.isinf(v):
        s = "-Inf" if v < 0 else "Inf"
    else:
        mant, exp = f"{abs(v):.{d}E}".split("E")
        e = int(exp)
        sign = "+" if e >= 0 else "-"
        s = mant + (f"E{sign}{abs(e):02d}" if abs(e) <= 99 else f"{sign}{abs(e):03d}")
        if math.copysign(1.0, v) < 0:
            s = "-" + s
    return "*" * w if len(s) > w else s.rjust(w)


def fortran_f(v: float, w: int, d: int) -> str:
    """Fw.d, right-justified in w."""
    s = f"{float(v):.{d}f}"
    if s.startswith("0.") and w < len(s):
        s = s[1:]
    return "*" * w if len(s) > w else s.rjust(w)


def npoints_1_to_np1(ncell: int, every: int) -> int:
    """mod_output_2d.f90:13-16."""
    return (ncell + 1 - 1) // every + 1


def _write_plane(filename: str, f2: np.ndarray, na: int, nb: int, every: int):
    """Rows b = nb+1 down to 1 (step every), values a = 1..na+1 (step every),
    as ES18.10 separated by one blank (the 1X after the last value moves the
    position only, so no record ends in a blank); the header is a
    list-directed write of the two point counts (written as amdflang's
    run-time writes it: one blank before each value)."""
    rows = [f" {npoints_1_to_np1(na, every)} {npoints_1_to_np1(nb, every)}"]
    for b in range(nb + 1, 0, -every):
        rows.append(" ".join(fortran_es(f2[b, a], 18, 10) for a in range(1, na + 2, every)))
    with open(filename, "w") as fh:
        fh.write("\n".join(rows) + "\n")


def _host(a) -> np.ndarray:
    return a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)


def write_plane_xy_scalar_2d(filename, f, n, every=1):
    _write_plane(filename, _host(f), int(n[0]), int(n[1]), every)


def write_plane_xz_scalar_2d(filename, f, n, every=1):
    _write_plane(filename, _host(f), int(n[0]), int(n[2]), every)


def write_plane_yz_scalar_2d(filename, f, n, every=1):
    _write_plane(filename, _host(f), int(n[1]), int(n[2]), every)


def write_plane_xy_scalar(filename, f, n, iz_plane, every=1):
    write_plane_xy_scalar_2d(filename, _host(f)[iz_plane], n, every)


def write_plane_xz_scalar(filename, f, n, iy_plane, every=1):
    write_plane_xz_scalar_2d(filename, _host(f)[:, iy_plane, :], n, every)


def write_plane_yz_scalar(filename, f, n, ix_plane, every=1):
    write_plane_yz_scalar_2d(filename, _host(f)[:, :, ix_plane], n, every)


def write_scalar_planes(f, n, ix_plane, iy_plane, iz_plane, every, prefix):
    """mod_output_2d.f90:82-97: <prefix>_xy.mco, _xz.mco, _yz.mco."""
    a = _host(f)
    write_plane_xy_scalar(f"{prefix}_xy.mco", a, n, iz_plane, every)
    write_plane_xz_scalar(f"{prefix}_xz.mco", a, n, iy_plane, every)
    write_plane_yz_scalar(f"{prefix}_yz.mco", a, n, ix_plane, every)


def write_density_planes(np_red, n, ptype, ix_plane, iy_plane, iz_plane, every, prefix):
    """mod_output_2d.f90:63-79 for species ptype (1-based) of np_red
    (ntype, n3+3, n2+3, n1+3)."""
    write_scalar_planes(_host(np_red)[ptype - 1], n, ix_plane, iy_plane, iz_plane, every, prefix)


def write_vector_component_planes(E4, n, comp, ix_plane, iy_plane, iz_plane, every, prefix):
    """mod_output_2d.f90:100-118: component comp (1-based) of the node field
    E4 (n3+3, n2+3, n1+3, 4)."""
    write_scalar_planes(_host(E4)[..., comp - 1], n, ix_plane, iy_plane, iz_plane, every, prefix)


# ---------------------------------------------------------------------------
# the driver's diagnostics
# ---------------------------------------------------------------------------

def _fsum(values) -> float:
    """Fortran SUM over an array section: left to right in element order."""
    s = 0.0
    for v in values:
        s = s + float(v)
    return s


class Diagnostics:
    """The modular driver's diagnostics for one Simulation.  Attach with
    sim.diagnostics = Diagnostics(sim, ...); Simulation calls the hooks.

    nsav: file and summary cadence; nb_step_averaging: moments cadence
    (mod_simParams.f90:140-144: 5 * nb_step_sort when a backup cadence is
    set, else nsav); planes: the 1-based node indices of the plot planes
    (defaults ix = n1/2+1, iz = n3/2+1, mod_simParams.f90:159-164); out_dir:
    where the .mco files go (None: no files)."""

    def __init__(self, sim, nsav: int, nb_step_averaging: int | None = None,
                 ix_plane: int | None = None, iz_plane: int | None = None,
                 out_dir: str | None = None, names=None):
        self.sim = sim
        n = sim.n
        self.n = n
        self.nsav = int(nsav)
        self.nb_step_averaging = int(nsav if nb_step_averaging is None else nb_step_averaging)
        self.ix_plane = int(n[0] // 2 + 1 if ix_plane is None else ix_plane)
        self.iz_plane = int(n[2] // 2 + 1 if iz_plane is None else iz_plane)
        self.iy_density = n[1] // 2 + 1
        self.iy_phi = n[1] // 2
        self.out_dir = out_dir
        self.names = list(names) if names is not None else [f"{i + 1}" for i in range(sim.ntype)]
        dev, nt = sim.dev, sim.ntype
        sxy, sxz, syz = plane_shapes(n)
        z = lambda *s: torch.zeros(s, dtype=torch.float64, device=dev)  # noqa: E731
        self.np_avg_xy, self.np_avg_xz, self.np_avg_yz = z(nt, *sxy), z(nt, *sxz), z(nt, *syz)
        self.phi_avg_xy, self.phi_avg_xz, self.phi_avg_yz = z(*sxy), z(*sxz), z(*syz)
        self.data_pavg_xy = z(nt, *sxy, 5)
        self.data_pavg_xz = z(nt, *sxz, 5)
        self.data_pavg_yz = z(nt, *syz, 5)
        self.cnt_avg = 0
        self.last = None

    # -- mod_state.f90:616-634 ------------------------------------------------
    def compute_plane_moments_local(self):
        s = self.sim
        for t in range(s.ntype):
            compute_particle_plane_moments_species(
                s.parts[t], s.n, s.h, s.case.dt, s.mass[t], self.ix_plane, self.iy_density,
                self.iz_plane, self.data_pavg_xy[t], self.data_pavg_xz[t], self.data_pavg_yz[t])

    # -- mod_state.f90:636-667 ------------------------------------------------
    def accumulate_2d_averages(self):
        s = self.sim
        npr = s.np_red
        self.np_avg_xy += npr[:, self.iz_plane, :, :]
        self.np_avg_xz += npr[:, :, self.iy_density, :]
        self.np_avg_yz += npr[:, :, :, self.ix_plane]
        self.phi_avg_xy += s.phi[self.iz_plane, :, :]
        self.phi_avg_xz += s.phi[:, self.iy_phi, :]
        self.phi_avg_yz += s.phi[:, :, self.ix_plane]
        self.cnt_avg += 1

    # -- mod_simulation.f90:224-240 -------------------------------------------
    def reset_2d_averages(self):
        for a in (self.np_avg_xy, self.np_avg_xz, self.np_avg_yz, self.phi_avg_xy,
                  self.phi_avg_xz, self.phi_avg_yz, self.data_pavg_xy, self.data_pavg_xz,
                  self.data_pavg_yz):
            a.zero_()
        self.cnt_avg = 0

    # -- hooks Simulation.advance_one_step calls -------------------------------
    def before_push(self, istep: int):
        if self.nb_step_averaging > 0 and istep % self.nb_step_averaging == 1:
            self.compute_plane_moments_local()
            self.accumulate_2d_averages()

    def after_step(self, istep: int):
        if istep % self.nsav == 1:
            if self.out_dir is not None:
                self.output_step(istep)
            self.reset_2d_averages()

    # -- mod_simulation.f90:243-319 -------------------------------------------
    def output_step(self, istep: int):
        n, d = self.n, self.out_dir
        os.makedirs(d, exist_ok=True)
        f = float(max(1, self.cnt_avg))
        for i in range(self.sim.ntype):
            pre = os.path.join(d, f"it{istep}_n{i + 1}")
            write_plane_xy_scalar_2d(pre + "_xy.mco", _host(self.np_avg_xy[i]) / f, n)
            write_plane_xz_scalar_2d(pre + "_xz.mco", _host(self.np_avg_xz[i]) / f, n)
            write_plane_yz_scalar_2d(pre + "_yz.mco", _host(self.np_avg_yz[i]) / f, n)
            pre = os.path.join(d, f"it{istep}_T{i + 1}")
            write_plane_xy_scalar_2d(pre + "_xy.mco", _host(self.data_pavg_xy[i, ..., TP_AVG]) / f, n)
            write_plane_xz_scalar_2d(pre + "_xz.mco", _host(self.data_pavg_xz[i, ..., TP_AVG]) / f, n)
            write_plane_yz_scalar_2d(pre + "_yz.mco", _host(self.data_pavg_yz[i, ..., TP_AVG]) / f, n)
        pre = os.path.join(d, f"it{istep}_phi")
        write_plane_xy_scalar_2d(pre + "_xy.mco", _host(self.phi_avg_xy) / f, n)
        write_plane_xz_scalar_2d(pre + "_xz.mco", _host(self.phi_avg_xz) / f, n)
        write_plane_yz_scalar_2d(pre + "_yz.mco", _host(self.phi_avg_yz) / f, n)
        for c, name in ((1, "Ex"), (2, "Ey"), (3, "Ez")):
            write_vector_component_planes(self.sim.E4, n, c, self.ix_plane, self.iy_density,
                                          self.iz_plane, 1, os.path.join(d, f"it{istep}_{name}"))

    # -- mod_simulation.f90:99-146 ---------------------------------------------
    def write_initial_diagnostics(self):
        n, d, s = self.n, self.out_dir, self.sim
        os.makedirs(d, exist_ok=True)
        for i in range(s.ntype):
            write_density_planes(s.np_red, n, i + 1, self.ix_plane, self.iy_density, self.iz_plane,
                                 1, os.path.join(d, f"n{i + 1}"))
        write_scalar_planes(s.phi, n, self.ix_plane, self.iy_phi, self.iz_plane, 1,
                            os.path.join(d, "phi1"))
        for c, name in ((1, "Ex"), (2, "Ey"), (3, "Ez")):
            write_vector_component_planes(s.E4, n, c, self.ix_plane, self.iy_density,
                                          self.iz_plane, 1, os.path.join(d, name))

    # -- mod_simulation.f90:574-694 ---------------------------------------------
    def power_diagnostics(self) -> dict:
        """The reductions of print_diagnostics over the accumulators since
        the last call: wall power and current (P_loss, p_mac, from the wall
        BC), absorbed heating power (P_heat), collision power (Pcoll, from
        MCC); injection is not part of this step and reads 0.  Resets the
        accumulators."""
        s = self.sim
        span = np.float64(float(self.nsav - 1) * s.case.dt)   # 0 when nsav = 1, as there
        P_loss = _host(s.P_loss)
        pm = _host(s.p_mac)                    # (ntype, slots, 2): charge, |charge| Ek
        Pcoll = _host(s.Pcoll) if getattr(s, "tables", None) is not None else np.zeros(1)
        q1 = _fsum(pm[0, :, 0])
        qi = _fsum(pm[1:, :, 0].reshape(-1))
        absi = _fsum(np.abs(pm[1:, :, 0]).reshape(-1))
        P_heat = _host(s.P_heat) if getattr(s, "P_heat", None) is not None else np.zeros(1)
        with np.errstate(divide="ignore", invalid="ignore"):
            out = {"Pwall": float(-np.float64(_fsum(P_loss)) / span),
                   "Pabs": float(np.float64(_fsum(P_heat)) / span),
                   "Pcoll": float(np.float64(_fsum(Pcoll)) / span), "Pinj": 0.0,
                   "Iw1": float(np.float64(q1) / span), "Iw2": float(np.float64(qi) / span)}
        out["Ekw1"] = _fsum(pm[0, :, 1]) / abs(q1) if abs(q1) > 0.0 else 0.0
        out["Ekw2"] = _fsum(pm[1:, :, 1].reshape(-1)) / absi if absi > 0.0 else 0.0
        s.P_loss.zero_()
        s.p_mac.zero_()
        if getattr(s, "P_heat", None) is not None:
            s.P_heat.zero_()
        if getattr(s, "tables", None) is not None:
            s.Pcoll.zero_()
        self.last = out
        return out

    def print_diagnostics(self, istep: int) -> str:
        """The text print_diagnostics writes (particle counts, power lines);
        returned, not printed."""
        s = self.sim
        v = self.power_diagnostics()
        counts = [p.n - h for p, h in zip(s.parts, s.lost_acc.tolist())]
        lines = ["", "", f" TIME STEP {istep} --> {fortran_f(s.case.dt * istep * 1.0e6, 8, 3)} us",
                 " -------------------------", " Particles diagnostics: "]
        lines += [f" npart {nm} = {c}" for nm, c in zip(self.names, counts)]
        lines += [" -------------------------", " Power diagnostics: ",
                  f" Pwall (W)  = {fortran_es(v['Pwall'], 10, 2)}",
                  f" Pabs  (W)  = {fortran_es(v['Pabs'], 10, 2)}",
                  f" Pcoll (W)  = {fortran_es(v['Pcoll'], 10, 2)}",
                  f" Pinj  (W)  = {fortran_es(v['Pinj'], 10, 2)}",
                  f" I_w  (A)  = {fortran_es(v['Iw1'], 10, 2)}{fortran_es(v['Iw2'], 10, 2)}",
                  f" Ek_w (eV) = {fortran_es(v['Ekw1'], 10, 2)}{fortran_es(v['Ekw2'], 10, 2)}",
                  " -------------------------"]
        return "\n".join(lines) + "\n"


__all__ = ["compute_particle_plane_moments_species", "Diagnostics", "fortran_es", "fortran_f",
           "npoints_1_to_np1", "write_plane_xy_scalar_2d", "write_plane_xz_scalar_2d",
           "write_plane_yz_scalar_2d", "write_plane_xy_scalar", "write_plane_xz_scalar",
           "write_plane_yz_scalar", "write_scalar_planes", "write_density_planes",
           "write_vector_component_planes", "QE"]
