"""Grid side of the step -- the call surface of modules/fields and the legacy
Poisson solver.

Grid arrays are FP64 device tensors of shape (n3+3, n2+3, n1+3) over the
reference's 0:n+2 node range (x fastest); E is (n3+3, n2+3, n1+3, 4) with
components (Ex, Ey, Ez, pad); bcnd is int32 of the node shape; np_thread and
np_red are (ntype, n3+3, n2+3, n1+3).
"""
from __future__ import annotations

import ctypes as C
import math

import torch

from . import _native
from .device import context, ptr
from .particles import ParticleSet

EPS0 = 8.854187817e-12


def build_kq(bcnd: torch.Tensor, n, kq: torch.Tensor | None = None) -> torch.Tensor:
    """kq = 2 where bcnd /= -1, else 1 (mod_charge_weights.f90:9-33)."""
    ctx = context(bcnd.device)
    kq = torch.empty(bcnd.shape, dtype=torch.float64, device=bcnd.device) if kq is None else kq
    g = _native.grid(n, (1.0, 1.0, 1.0))
    ctx.check(ctx.lib.lepic_build_kq(ctx.h, C.byref(g), ptr(bcnd), ptr(kq)))
    return kq


def clear_np_thread(np_thread: torch.Tensor):
    """np_thread = 0 (mod_chargeDeposition.f90:16-22)."""
    ctx = context(np_thread.device)
    ctx.check(ctx.lib.lepic_fill(ctx.h, ptr(np_thread), np_thread.numel(), 0.0))


def deposit_particle_set_to_np_thread(part: ParticleSet, n, h, kq: torch.Tensor, Nm_species: float,
                                      np_local: torch.Tensor):
    """Trilinear charge deposition (mod_chargeDeposition.f90:25-133)."""
    ctx = context(part.device)
    g = _native.grid(n, h)
    p = part.pset()
    ctx.check(ctx.lib.lepic_deposit(ctx.h, C.byref(g), C.byref(p), ptr(kq), float(Nm_species),
                                    ptr(np_local)))


def fixed_point_bits(n, h, Nm_species: float) -> int:
    """Fraction bits for deposit_fixed: a full weight |Nm/(h1 h2 h3)| maps
    to at most 2^44 units, leaving a node's int64 total room for 2^19
    particles."""
    k1 = abs(float(Nm_species) / (h[0] * h[1] * h[2]))
    if k1 == 0.0:
        return 0        # nothing is deposited
    return 44 - math.ceil(math.log2(k1))


def fixed_point_bits32(n, h, Nm_species: float) -> int:
    """Fraction bits for the FP32 particle phase's deposit: a full weight
    |Nm/(h1 h2 h3)| maps to at most 2^21 units (its FP32 split into 8 node
    parts is exact up to 2^22, and one block's sums at a node stay within
    uint32)."""
    k1 = abs(float(Nm_species) / (h[0] * h[1] * h[2]))
    if k1 == 0.0:
        return 0
    return 21 - math.ceil(math.log2(k1))


def deposit_particle_set_to_np_thread_fixed(part: ParticleSet, n, h, kq: torch.Tensor,
                                            Nm_species: float, np_local: torch.Tensor,
                                            frac_bits: int | None = None):
    """deposit_particle_set_to_np_thread (mod_chargeDeposition.f90:25-133)
    in fixed point: each particle's full weight, rint(|Nm/(h1 h2 h3)| *
    2^frac_bits) units, is split into 8 integer node parts that sum to it
    exactly and summed as int64, so the density does not depend on particle
    order and the charge on the grid is exactly the number of deposited
    particles times one full weight."""
    if frac_bits is None:
        frac_bits = fixed_point_bits(n, h, Nm_species)
    ctx = context(part.device)
    g = _native.grid(n, h)
    p = part.pset()
    ctx.check(ctx.lib.lepic_deposit_fixed(ctx.h, C.byref(g), C.byref(p), ptr(kq),
                                          float(Nm_species), int(frac_bits), ptr(np_local)))
    return frac_bits


def reduce_species_density(n, bcnd: torch.Tensor, np_thread: torch.Tensor, ntype: int,
                           np_red: torch.Tensor):
    """Periodic stitching (in place on np_thread) and the sum over the one
    virtual thread onto nodes 1..n+1 (mod_density.f90:31-73)."""
    ctx = context(np_thread.device)
    g = _native.grid(n, (1.0, 1.0, 1.0))
    ctx.check(ctx.lib.lepic_reduce_density(ctx.h, C.byref(g), ptr(bcnd), int(ntype), ptr(np_thread),
                                           ptr(np_red)))


def build_rho_from_np(n, np_red: torch.Tensor, charge, ntype: int, rho: torch.Tensor):
    """rho = -sum_s q_s np_s on nodes 1..n+1 (mod_density.f90:76-107)."""
    ctx = context(np_red.device)
    g = _native.grid(n, (1.0, 1.0, 1.0))
    q = (C.c_double * _native.MAX_SPECIES)(*[float(v) for v in charge])
    ctx.check(ctx.lib.lepic_build_rho(ctx.h, C.byref(g), int(ntype), ptr(np_red), q, ptr(rho)))


def solve_poisson_legacy(phi: torch.Tensor, bcnd: torch.Tensor, rhs: torch.Tensor, h, n,
                         ncycl: int, eps: float, omega: float, ng: int, flag_pbc: int,
                         flag_nmn: int, eps0: float = EPS0, graph: bool = True) -> dict:
    """V-cycle multigrid / red-black SOR, phi updated in place from its
    current value (mod_PoissonSolver_legacy.f90:27-101 -> pdesolver.f90).
    Returns the V-cycle count, equivalent sweeps and final L1/alpha residual.
    graph: replay each V-cycle's launches from one captured HIP graph."""
    ctx = context(phi.device)
    g = _native.grid(n, h)
    mp = _native.MGParams()
    mp.flag_pbc, mp.flag_nmn, mp.eps0 = int(flag_pbc), int(flag_nmn), float(eps0)
    mp.ncycl, mp.eps, mp.omega, mp.ng = int(ncycl), float(eps), float(omega), int(ng)
    mp.graph = 1 if graph else 0
    st = _native.MGStats()
    ctx.check(ctx.lib.lepic_poisson_solve(ctx.h, C.byref(g), ptr(phi), ptr(rhs), ptr(bcnd),
                                          C.byref(mp), C.byref(st)))
    return {"cycles": st.cycles, "ktot": st.ktot, "res": st.res,
            "sweeps": [st.sweeps[i] for i in range(int(ng))]}


def solve_poisson_mixed(phi: torch.Tensor, bcnd: torch.Tensor, rhs: torch.Tensor, h, n,
                        ncycl: int, eps: float, omega: float, ng: int, flag_pbc: int,
                        flag_nmn: int, eps0: float = EPS0, inner_cycles: int = 4, nu: int = 2,
                        nu_coarsest: int = 64, omega_smooth: float = 1.15,
                        graph: bool = True) -> dict:
    """The Poisson solve in its precision form: FP32 red-black smoothing
    V-cycles (inner_cycles per step, nu sweeps per level and leg at
    omega_smooth) inside an FP64 iterative refinement, stopped as
    solve_poisson_legacy stops (FP64 L1 residual / alpha < eps).  omega (the
    reference's over-relaxation) is not used by the FP32 smoother.  Returns
    the refinement steps ("cycles"), equivalent fine sweeps, the residual
    and the FP32 sweeps per level."""
    ctx = context(phi.device)
    g = _native.grid(n, h)
    mp = _native.MGParams()
    mp.flag_pbc, mp.flag_nmn, mp.eps0 = int(flag_pbc), int(flag_nmn), float(eps0)
    mp.ncycl, mp.eps, mp.omega, mp.ng = int(ncycl), float(eps), float(omega), int(ng)
    mp.graph = 1 if graph else 0
    xp = _native.MGMixParams()
    xp.inner_cycles, xp.nu, xp.nu_coarsest = int(inner_cycles), int(nu), int(nu_coarsest)
    xp.omega = float(omega_smooth)
    st = _native.MGStats()
    ctx.check(ctx.lib.lepic_poisson_solve_mixed(ctx.h, C.byref(g), ptr(phi), ptr(rhs), ptr(bcnd),
                                                C.byref(mp), C.byref(xp), C.byref(st)))
    return {"cycles": st.cycles, "ktot": st.ktot, "res": st.res,
            "sweeps": [st.sweeps[i] for i in range(int(ng))]}


def solve_poisson_pcg(phi: torch.Tensor, bcnd: torch.Tensor, rhs: torch.Tensor, h, n,
                      ncycl: int, eps: float, omega: float, ng: int, flag_pbc: int = 0,
                      flag_nmn: int = 0, eps0: float = EPS0, inner_cycles: int = 1, nu: int = 2,
                      nu_coarsest: int = 16, omega_smooth: float = 1.15,
                      graph: bool = True, precond_fp64: bool = False) -> dict:
    """The Poisson solve by FP64 conjugate gradients preconditioned with the
    mixed solve's fixed-sweep V-cycles (inner_cycles per iteration), in FP32
    or, with precond_fp64, in FP64 (the whole solve FP64), 16 sweeps on the
    coarsest level (the CG corrects what a shorter coarse solve leaves),
    stopped as solve_poisson_legacy stops; plasma and electrode labels only, no
    periodic or Neumann walls.  Returns the iterations ("cycles"), equivalent
    fine sweeps, the residual and the preconditioner's sweeps per level."""
    if flag_pbc or flag_nmn:
        raise ValueError("the conjugate-gradient solve takes no periodic or Neumann walls")
    ctx = context(phi.device)
    g = _native.grid(n, h)
    mp = _native.MGParams()
    mp.flag_pbc, mp.flag_nmn, mp.eps0 = 0, 0, float(eps0)
    mp.ncycl, mp.eps, mp.omega, mp.ng = int(ncycl), float(eps), float(omega), int(ng)
    mp.graph = 1 if graph else 0
    xp = _native.MGMixParams()
    xp.inner_cycles, xp.nu, xp.nu_coarsest = int(inner_cycles), int(nu), int(nu_coarsest)
    xp.omega = float(omega_smooth)
    xp.precond_fp64 = 1 if precond_fp64 else 0
    st = _native.MGStats()
    ctx.check(ctx.lib.lepic_poisson_solve_pcg(ctx.h, C.byref(g), ptr(phi), ptr(rhs), ptr(bcnd),
                                              C.byref(mp), C.byref(xp), C.byref(st)))
    return {"cycles": st.cycles, "ktot": st.ktot, "res": st.res,
            "sweeps": [st.sweeps[i] for i in range(int(ng))]}


def apply_dielectric_bc_to_phi(n, h, bcnd: torch.Tensor, dtype: torch.Tensor, ngrid: int,
                               sum_q_xz: torch.Tensor, sum_q_yz: torch.Tensor, Ca_cell: float,
                               phi: torch.Tensor):
    """apply_dielectric_bc_to_phi (mod_state.f90:378-430): dielectric nodes
    take each species' collected surface charge.  sum_q_xz (ntype, n3+3,
    n1+3), sum_q_yz (ntype, n3+3, n2+3, 2), dtype int32 over labels 0..ngrid;
    Ca_cell = Ca h1 h3."""
    ntype = int(sum_q_xz.shape[0])
    ctx = context(phi.device)
    ctx.check(ctx.lib.lepic_dielectric_phi(ctx.h, C.byref(_native.grid(n, h)), ptr(bcnd), ptr(dtype),
                                           int(ngrid), ntype, ptr(sum_q_xz), ptr(sum_q_yz),
                                           float(Ca_cell), ptr(phi)))


def calc_Efield_modular(n, h, phi: torch.Tensor, E4: torch.Tensor, bcnd: torch.Tensor):
    """E = -grad(phi) with the reference's wall extrapolation and periodic
    ghost copies (mod_electricField.f90:9-164)."""
    ctx = context(phi.device)
    g = _native.grid(n, h)
    ctx.check(ctx.lib.lepic_efield(ctx.h, C.byref(g), ptr(phi), ptr(bcnd), ptr(E4)))
