"""Host-side set-up of the PIC-MCC step: constants, grids, boundary labels,
species parameters and the reaction tables MCC consumes.

This is one-time host work (not on the timed path).  It restates the pieces
of the reference's set-up that define the step's inputs:

* physical constants            -- modules/core/mod_constants.f90:7-12,
                                   modules/particles/mod_reactions.f90:107-110
* vt0, Nm, dt                   -- modules/simulation/mod_simParams.f90:98-130
* kq (charge weights)           -- modules/fields/mod_charge_weights.f90:9-33
* common energy grid, sig,
  sig_list, sigv_mx, p_ncol     -- modules/particles/mod_reactions.f90:414-558
* cadences and nu_uplim         -- modules/simulation/mod_simParams.f90:132-156

Grid arrays are numpy arrays of shape (n3+3, n2+3, n1+3): node (i,j,k) of the
reference's 0:n+2 index space is element [k, j, i] (x fastest in memory).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

# mod_constants.f90 / mod_reactions.f90
QE = 1.60217646e-19
ME = 9.10938188e-31
AMU = 1.66053886e-27
EPS0 = 8.854187817e-12
KB = 1.380649e-23
MAX_COL = 128
NU_UPLIM_ELECTRON = 5.0e8      # mod_simParams.f90:155
NU_UPLIM_OTHER = 1.0e7         # mod_simParams.f90:156


def node_shape(n):
    return (int(n[2]) + 3, int(n[1]) + 3, int(n[0]) + 3)


def grounded_box_bcnd(n, label: int = 1) -> np.ndarray:
    """Box whose six faces are one electrode: label on nodes {0,1,n+1,n+2} of
    every dimension, plasma (-1) on nodes 2..n."""
    b = np.full(node_shape(n), -1, np.int32)
    for axis, nd in zip((2, 1, 0), n):
        sl = [slice(None)] * 3
        for idx in (0, 1, nd + 1, nd + 2):
            sl[axis] = idx
            b[tuple(sl)] = label
    return b


def biased_wall_bcnd(n, wall_label: int = 2) -> np.ndarray:
    """Grounded box (label 1) whose x = L face (nodes n1+1, n1+2) carries its
    own electrode label."""
    b = grounded_box_bcnd(n, 1)
    b[:, :, n[0] + 1:] = wall_label
    return b


def build_kq(bcnd: np.ndarray) -> np.ndarray:
    """kq = 2 on every node with bcnd /= -1, else 1 (mod_charge_weights.f90:22-30)."""
    return np.where(bcnd != -1, 2.0, 1.0)


def omega_pe(n0: float) -> float:
    return math.sqrt(n0 * QE * QE / (EPS0 * ME))


def thermal_speed(T_eV: float, mass: float) -> float:
    """vt0 = sqrt(2 q T / |m|) (mod_simParams.f90:112-116)."""
    return math.sqrt(2.0 * QE * T_eV / abs(mass))


@dataclasses.dataclass
class Species:
    name: str
    mass: float          # kg, neutrals negative as in the reference
    charge: float        # C
    T_eV: float
    tracked: bool = True


@dataclasses.dataclass
class Case:
    """One run configuration: grid, boundary, species and step parameters."""
    name: str
    n: tuple
    h: tuple
    bcnd: np.ndarray
    ngrid: int
    electrode_V: dict
    species: list
    n0: float
    np_cell: int
    dt: float
    steps: int
    eps: float = 1e-8
    omega: float = 1.4
    ng: int = 5
    ncycl: int = 20000
    flag_pbc: int = 0
    flag_nmn: int = 0
    B_uniform: tuple | None = None
    B_map: np.ndarray | None = None
    B_grid: tuple | None = None          # (nB, hB) when B_map lives on its own grid
    ions_magnetized: bool = True
    ngas: float = 0.0
    nb_step_sort: int = 10
    nb_step_collisions: int = 10
    chemistry: dict | None = None
    phi_init: np.ndarray | None = None   # initial potential (a deck's generate_boundary u)
    ni0: list | None = None              # loader acceptance n/n0 per tracked species
    x_load: float = 0.0                  # loading region x < x_load (m); <= 0: whole box
    heating: dict | None = None          # electron heating (a deck's; None: no heating)
    flag_die: int = 0                    # dielectric surface charging
    dtype: np.ndarray | None = None      # wall type per label 0..ngrid (flag_die)
    Ca: float = 0.0                      # dielectric capacitance per area (F/m^2)
    # loss plane of a differentiable run: (axis, 1-based node plane, sign,
    # area in m^2 or None for the whole face); None: x = (n1 - 1) h1, sign +1
    flux_plane: tuple | None = None

    @property
    def tracked(self):
        return [s for s in self.species if s.tracked]

    @property
    def vt0(self):
        return [thermal_speed(s.T_eV, s.mass) for s in self.species]

    @property
    def Nm(self):
        """Macro-particle weight (mod_simParams.f90:118-122), with the
        reference's order of operations: n * h1 * h2 * h3 / |np_cell|."""
        return [(self.n0 if s.charge != 0.0 else self.ngas) * self.h[0] * self.h[1] * self.h[2]
                / abs(self.np_cell) for s in self.species]

    def phi0(self) -> np.ndarray:
        """Initial potential: the deck's generate_boundary potential when the
        case has one, else electrode voltages on their labelled nodes."""
        if self.phi_init is not None:
            return np.array(self.phi_init, np.float64)
        phi = np.zeros(node_shape(self.n))
        for lab, v in self.electrode_V.items():
            phi[self.bcnd == lab] = v
        return phi

    def kq(self):
        return build_kq(self.bcnd)


def argon_species(Te=3.0, Ti=0.03, Tgas_K=300.0):
    return [Species("e", ME, -QE, Te), Species("Ar+", 39.95 * AMU, QE, Ti),
            Species("Ar", -39.95 * AMU, 0.0, Tgas_K * KB / QE, tracked=False)]


def cfg1_plain_box(n: int = 32, side: float = 0.01, np_cell: int = 20, steps: int = 100,
                   ng: int = 5) -> Case:
    """Plain PIC: grounded cubic box, e- + Ar+, no B, no collisions,
    dt = 0.1/omega_pe, MG to 1e-8 (reference L1/alpha residual)."""
    nn = (n, n, n)
    h = (side / n,) * 3
    n0 = 1e15
    sp = argon_species()[:2]
    return Case("plain-box", nn, h, grounded_box_bcnd(nn), 1, {1: 0.0}, sp, n0, np_cell,
                0.1 / omega_pe(n0), steps, ng=ng, nb_step_collisions=0)


def cfg2_pic_mcc(n: int = 128, np_cell: int = 64, steps: int = 1000, ng: int = 7,
                 pressure_Pa: float = 1.0, Bz: float = 0.01, deck=None) -> Case:
    """PIC-MCC at scale: h = 3.125e-4 m (4 cm box at 128^3), uniform Bz,
    argon background, analytic cross sections, MCC every 10 steps, sort every 20.
    deck: a chemistry deck in the reference's format (path or text) whose
    species and reactions replace the analytic argon set."""
    nn = (n, n, n)
    h = (3.125e-4,) * 3
    n0 = 1e15
    sp = argon_species()
    ngas = pressure_Pa / (KB * 300.0)
    if deck is not None:
        from . import chemistry
        chem = chemistry.read_reactions(deck, ngas, Te=sp[0].T_eV)
        c = Case("pic-mcc", nn, h, grounded_box_bcnd(nn), 1, {1: 0.0}, chem.species, n0,
                 np_cell, 0.1 / omega_pe(n0), steps, ng=ng, B_uniform=(0.0, 0.0, Bz), ngas=ngas,
                 nb_step_sort=20, nb_step_collisions=10)
        c.chemistry = chemistry.build_tables(chem, ngas)
        return c
    c = Case("pic-mcc", nn, h, grounded_box_bcnd(nn), 1, {1: 0.0}, sp, n0, np_cell,
             0.1 / omega_pe(n0), steps, ng=ng, B_uniform=(0.0, 0.0, Bz), ngas=ngas,
             nb_step_sort=20, nb_step_collisions=10)
    c.chemistry = argon_chemistry(sp, ngas)
    return c


def cfg4_differentiable(n: int = 64, np_cell: int = 8, Vb: float = 20.0, B0: float = 0.005,
                        steps: int = 200, ng: int | None = None, mcc_every: int = 10,
                        pressure_Pa: float = 1.0) -> Case:
    """Differentiable PIC: n^3 cells (h = 3.125e-4 m), e- and Ar+ at np_cell
    particles per cell and species, a grounded box whose x = L face is an
    electrode of its own (label 2) biased at Vb, uniform Bz = B0 acting on
    both species, dt = 0.1/omega_pe, MG to 1e-10; the analytic argon MCC of
    the PIC-MCC case every mcc_every steps (0: no collisions) on a background
    at pressure_Pa and 300 K.  The derivatives are taken with respect to Vb
    and B0."""
    nn = (n, n, n)
    h = (3.125e-4,) * 3
    n0 = 1e15
    if ng is None:
        ng = max(2, int(round(math.log2(n))))
    if mcc_every <= 0:
        sp = argon_species()[:2]
        return Case("differentiable", nn, h, biased_wall_bcnd(nn, 2), 2, {1: 0.0, 2: Vb}, sp,
                    n0, np_cell, 0.1 / omega_pe(n0), steps, eps=1e-10, ng=ng,
                    B_uniform=(0.0, 0.0, B0), nb_step_collisions=0)
    sp = argon_species()
    ngas = pressure_Pa / (KB * 300.0)
    c = Case("differentiable", nn, h, biased_wall_bcnd(nn, 2), 2, {1: 0.0, 2: Vb}, sp, n0,
             np_cell, 0.1 / omega_pe(n0), steps, eps=1e-10, ng=ng, B_uniform=(0.0, 0.0, B0),
             ngas=ngas, nb_step_collisions=int(mcc_every))
    c.chemistry = argon_chemistry(sp, ngas)
    return c


def loadable_cells(bcnd: np.ndarray, n) -> int:
    """Cells with a plasma (-1) node among their 8 corners: the cells the
    loaders fill (mod_particle_loader.f90:75-91)."""
    pl = np.asarray(bcnd) == -1
    n1, n2, n3 = (int(v) for v in n)
    load = np.zeros((n3, n2, n1), bool)
    for dk in (0, 1):
        for dj in (0, 1):
            for di in (0, 1):
                load |= pl[1 + dk:n3 + 1 + dk, 1 + dj:n2 + 1 + dj, 1 + di:n1 + 1 + di]
    return int(load.sum())


def annular_channel_bcnd(n, h, r_in: float = 0.3, r_out: float = 0.45, anode_label: int = 2):
    """A grounded box (label 1) holding a cylindrical annular channel along z
    about the box's centre line: plasma (-1) on the interior nodes whose
    distance r from the axis satisfies r_in L <= r <= r_out L (L the box
    width along x), label 1 elsewhere; the channel's cross-section on the
    z = 0 face (node planes 0 and 1) is the anode (anode_label)."""
    b = grounded_box_bcnd(n, 1)
    L = n[0] * h[0]
    x = (np.arange(n[0] + 3) - 1) * h[0] - 0.5 * L
    y = (np.arange(n[1] + 3) - 1) * h[1] - 0.5 * n[1] * h[1]
    r = np.sqrt(x[None, :] ** 2 + y[:, None] ** 2)          # (j, i)
    inside = (r >= r_in * L) & (r <= r_out * L)
    b[:, ~inside] = 1
    b[0:2][:, inside] = anode_label
    return b


def radial_bfield_map(n, h, B0: float = 0.02, z_peak: float = 0.7, width: float = 0.1):
    """Radial field B = Br(z) r_hat about the box's centre line, as a map of 4
    values per node (Bx, By, Bz, |B|) on the field grid: Br(z) = B0 exp(-(z -
    z_peak L)^2 / (2 (width L)^2)), L the box length along z (the Gaussian
    profile form of gaussian_Bfield, mod_magneticField.f90:410-428); zero on
    the axis."""
    Lx, Ly, Lz = (n[d] * h[d] for d in range(3))
    x = (np.arange(n[0] + 3) - 1) * h[0] - 0.5 * Lx
    y = (np.arange(n[1] + 3) - 1) * h[1] - 0.5 * Ly
    z = (np.arange(n[2] + 3) - 1) * h[2]
    br = B0 * np.exp(-(z - z_peak * Lz) ** 2 / (2.0 * (width * Lz) ** 2))      # (k,)
    r = np.sqrt(x[None, :] ** 2 + y[:, None] ** 2)                             # (j, i)
    with np.errstate(invalid="ignore", divide="ignore"):
        cx = np.where(r > 0.0, x[None, :] / r, 0.0)
        cy = np.where(r > 0.0, y[:, None] / r, 0.0)
    B4 = np.zeros(node_shape(n) + (4,))
    B4[..., 0] = br[:, None, None] * cx[None]
    B4[..., 1] = br[:, None, None] * cy[None]
    B4[..., 3] = np.sqrt(B4[..., 0] ** 2 + B4[..., 1] ** 2 + B4[..., 2] ** 2)
    return B4


def cfg5_combined(n: int = 256, np_cell: int = 32, steps: int = 500, V_anode: float = 300.0,
                  B0: float = 0.02, pressure_Pa: float = 1.0, ng: int | None = None) -> Case:
    """Combined large run: n^3 cells (h = 3.125e-4 m), the annular channel of
    annular_channel_bcnd (inner radius 0.3, outer 0.45 of the box width) with
    the anode at z = 0 biased at V_anode, the radial Br(z) map of
    radial_bfield_map (peak B0 at z = 0.7 L, width 0.1 L) acting on both
    species with the |B| > 1e-5 gate, e- and Ar+ at np_cell particles per
    loadable cell and species, the analytic argon MCC of the PIC-MCC case
    every 10 steps, sort every 20, dt = 0.1/omega_pe, Poisson to 1e-10."""
    nn = (n, n, n)
    h = (3.125e-4,) * 3
    n0 = 1e15
    if ng is None:
        ng = max(2, int(round(math.log2(n))))
    sp = argon_species()
    ngas = pressure_Pa / (KB * 300.0)
    c = Case("combined", nn, h, annular_channel_bcnd(nn, h), 2, {1: 0.0, 2: V_anode}, sp, n0,
             np_cell, 0.1 / omega_pe(n0), steps, eps=1e-10, ng=ng, ngas=ngas,
             nb_step_sort=20, nb_step_collisions=10)
    c.B_map = radial_bfield_map(nn, h, B0)
    c.chemistry = argon_chemistry(sp, ngas)
    # the differentiable run's loss: the ion current (A) into the anode,
    # through the node plane z = h3 one cell in front of it
    c.flux_plane = (2, 2, -1.0, 1.0)
    return c


def cfg3_poisson(n: int = 256, seed: int = 3, h: float = 3.125e-4, nblobs: int = 64):
    """Mixed-precision Poisson problem: an n^3-cell box whose six faces are
    electrodes 1..6 (painted x=0, x=L, y=0, y=L, z=0, z=L in that order, so a
    later face owns the shared edges and corners, as later segments of
    generate_boundary do) at voltages drawn U[-100, 100] V; rho a sum of
    nblobs Gaussian blobs (centres uniform in the box, sigma U[2, 10] cells,
    amplitudes N(0, 1e-4) C/m^3).  Returns (n, h, bcnd, phi0, rhs) with
    phi0 the electrode voltages (0 elsewhere) and rhs = -rho on plasma nodes
    (the sign build_rho_from_np gives, mod_density.f90:99)."""
    rng = np.random.default_rng(seed)
    nn = (n, n, n)
    hh = (h, h, h)
    b = np.full(node_shape(nn), -1, np.int32)
    V = rng.uniform(-100.0, 100.0, 6)
    label = 0
    for axis in (2, 1, 0):            # numpy axes of x, y, z
        for idx in ((0, 1), (n + 1, n + 2)):
            label += 1
            sl = [slice(None)] * 3
            sl[axis] = list(idx)
            b[tuple(sl)] = label
    phi0 = np.zeros(b.shape)
    for lab in range(1, 7):
        phi0[b == lab] = V[lab - 1]
    c = rng.uniform(0.0, n * h, (nblobs, 3))
    sig = rng.uniform(2.0, 10.0, nblobs) * h
    A = rng.normal(0.0, 1e-4, nblobs)
    x = (np.arange(n + 3) - 1) * h
    rho = np.zeros(b.shape)
    for k in range(nblobs):
        gx = np.exp(-(x - c[k, 0]) ** 2 / (2 * sig[k] ** 2))
        gy = np.exp(-(x - c[k, 1]) ** 2 / (2 * sig[k] ** 2))
        gz = np.exp(-(x - c[k, 2]) ** 2 / (2 * sig[k] ** 2))
        rho += A[k] * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    rhs = np.where(b == -1, -rho, 0.0)
    return nn, hh, b, phi0, rhs


# ---------------------------------------------------------------------------
# Reaction tables (mod_reactions.f90:414-558 "ordering")
# ---------------------------------------------------------------------------

S0_EXCITATION = 3.6e-18   # eV^2 m^2: peak ~1.0e-20 m^2 at e*11.5 eV
S0_IONIZATION = 1.9e-17   # eV^2 m^2: peak ~2.8e-20 m^2 at e*15.8 eV


def lotz_like(E: np.ndarray, s0: float, Eth: float) -> np.ndarray:
    """sigma(E) = s0 ln(E/Eth) / (E Eth) above threshold, 0 below."""
    out = np.zeros_like(E)
    m = E > Eth
    out[m] = s0 * np.log(E[m] / Eth) / (E[m] * Eth)
    return out


def argon_energy_grid(npts: int = 200, extra=(11.5, 15.8)) -> np.ndarray:
    grid = np.concatenate([[0.0], np.logspace(-2.0, 3.0, npts), np.asarray(extra, float)])
    return np.unique(grid)


def argon_chemistry(species, ngas: float, npts: int = 200) -> dict:
    """Synthetic argon set: e+Ar elastic (1e-19 m^2), excitation (11.5 eV) and
    ionization (15.8 eV) as s0 ln(E/Eth)/(E Eth), Ar+ + Ar charge exchange
    (5e-19 m^2).  Species indices are 0-based: e=0, Ar+=1, Ar=2."""
    E = argon_energy_grid(npts)
    reactions = [
        # (reactants, byproducts, Eth, sigma(E))
        ((0, 2), (0, 2), 0.0, np.full_like(E, 1e-19)),
        ((0, 2), (0, 2), 11.5, lotz_like(E, S0_EXCITATION, 11.5)),
        ((0, 2), (0, 1, 0), 15.8, lotz_like(E, S0_IONIZATION, 15.8)),
        ((1, 2), (1, 2), 0.0, np.full_like(E, 5e-19)),
    ]
    tables = [(E.copy(), r[3]) for r in reactions]
    return build_reaction_tables(species, ngas, reactions, tables)


def build_reaction_tables(species, ngas, reactions, tables, neutral_density=None,
                          kinds=None) -> dict:
    """Common energy grid, interpolated sig, sig_list, p_ncol and sigv_mx as
    mod_reactions.f90:414-558 builds them.  neutral_density: per neutral
    species (default ngas each); kinds: (reaction type, electrons among the
    products) per reaction, stored after the species in col_info as the
    reference's parser does (mod_reactions.f90:255-273)."""
    ntot = len(species)
    ntrk = sum(1 for s in species if s.tracked)
    ncol = len(reactions)
    merged = np.unique(np.concatenate([t[0] for t in tables]))   # sort + drop duplicates
    npt = merged.size
    sig = np.zeros((ncol, npt))
    for icol, (Ek, sg) in enumerate(tables):
        if Ek.size < 2:
            continue
        sig_p = 0.0
        for ipt in range(npt):
            Ekp = merged[ipt]
            for j in range(Ek.size - 1):
                EkL, EkR, sL, sR = Ek[j], Ek[j + 1], sg[j], sg[j + 1]
                if Ekp < EkL:
                    sig_p = sL
                    break
                if EkL <= Ekp <= EkR:
                    sig_p = sL + (Ekp - EkL) * (sR - sL) / (EkR - EkL)
                    break
            sig[icol, ipt] = sig_p
    col_info = np.zeros((ncol, 10), np.int32)
    sig_Eth = np.zeros(ncol)
    sig_list = np.zeros((ntot, MAX_COL), np.int32)
    p_ncol = np.zeros(ntot, np.int32)
    sigv_mx = np.zeros((ntot, ncol))
    masses = np.array([s.mass for s in species])
    for icol, (re, by, Eth, _) in enumerate(reactions):
        col_info[icol, 0] = len(re)
        col_info[icol, 1] = len(by)
        col_info[icol, 2:2 + len(re)] = re
        col_info[icol, 2 + len(re):2 + len(re) + len(by)] = by
        if kinds is not None and 2 + len(re) + len(by) + 1 < 10:
            col_info[icol, 2 + len(re) + len(by)] = kinds[icol][0]
            col_info[icol, 3 + len(re) + len(by)] = kinds[icol][1]
        sig_Eth[icol] = Eth
        r1, r2 = re[0], re[1]
        sig_list[r1, p_ncol[r1]] = icol
        p_ncol[r1] += 1
        mu = abs(masses[r1]) * abs(masses[r2]) / (abs(masses[r1]) + abs(masses[r2]))
        for ipt in range(npt):
            vr = math.sqrt(2.0 * merged[ipt] * QE / mu)
            if merged[ipt] >= 1.0:
                sigv_mx[r1, icol] = max(sigv_mx[r1, icol], vr * sig[icol, ipt])
    nu_uplim = np.array([NU_UPLIM_ELECTRON] + [NU_UPLIM_OTHER] * (ntot - 1))
    vt0 = np.array([thermal_speed(s.T_eV, s.mass) for s in species])
    return {
        "ntype_tracked": ntrk, "ntype_total": ntot, "mass": masses, "vt0": vt0,
        "neutral_density": np.array(neutral_density if neutral_density is not None
                                    else [ngas for s in species if not s.tracked], float),
        "nu_uplim": nu_uplim, "p_ncol": p_ncol, "sig_list": sig_list.reshape(-1),
        "ncol": ncol, "npt": npt, "sig_Er": merged, "sig": sig.reshape(-1), "sig_Eth": sig_Eth,
        "col_info": col_info.reshape(-1), "sigv_mx": sigv_mx.reshape(-1),
    }
