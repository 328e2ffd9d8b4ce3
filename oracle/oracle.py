"""ctypes front end of the CPU oracle (oracle/lepic_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  The product package never imports this module.

Arrays follow the conventions of lepic_oracle.h: grid arrays are flattened
x-fastest over nodes 0..n+2 (numpy shape (n3+3, n2+3, n1+3)), E and B maps
carry 4 doubles per node, cell ids and particle indices are 0-based.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liblepic_oracle.so")

MAX_COL = 128


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, strict IEEE)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


class Grid(C.Structure):
    _fields_ = [("n", C.c_int32 * 3), ("h", C.c_double * 3)]


class PSet(C.Structure):
    _fields_ = [
        ("x", C.POINTER(C.c_double)), ("y", C.POINTER(C.c_double)), ("z", C.POINTER(C.c_double)),
        ("vx", C.POINTER(C.c_double)), ("vy", C.POINTER(C.c_double)), ("vz", C.POINTER(C.c_double)),
        ("cell", C.POINTER(C.c_int32)), ("dead", C.POINTER(C.c_int8)),
        ("n", C.c_int64), ("cap", C.c_int64),
    ]


class MGStats(C.Structure):
    _fields_ = [("cycles", C.c_int32), ("ktot", C.c_double), ("res", C.c_double),
                ("sweeps", C.c_int64 * 12)]


class Chem(C.Structure):
    _fields_ = [
        ("ntype_tracked", C.c_int32), ("ntype_total", C.c_int32),
        ("mass", C.POINTER(C.c_double)), ("vt0", C.POINTER(C.c_double)),
        ("Nm", C.POINTER(C.c_double)), ("neutral_density", C.POINTER(C.c_double)),
        ("nu_uplim", C.POINTER(C.c_double)), ("p_ncol", C.POINTER(C.c_int32)),
        ("sig_list", C.POINTER(C.c_int32)), ("ncol", C.c_int32), ("npt", C.c_int32),
        ("sig_Er", C.POINTER(C.c_double)), ("sig", C.POINTER(C.c_double)),
        ("sig_Eth", C.POINTER(C.c_double)), ("col_info", C.POINTER(C.c_int32)),
        ("sigv_mx", C.POINTER(C.c_double)),
    ]


class Heating(C.Structure):
    _fields_ = [("nudt", C.c_double), ("Pabs", C.c_double), ("nu_h", C.c_double),
                ("xl_pow", C.c_double), ("xr_pow", C.c_double),
                ("flag_circxh", C.c_int32), ("flag_ahp", C.c_int32),
                ("R_ahp_moments", C.c_double), ("R_ahp", C.c_double),
                ("ymax", C.c_double), ("zmax", C.c_double), ("Nm", C.c_double),
                ("mass", C.c_double), ("vt0", C.c_double)]


class MCCStats(C.Structure):
    _fields_ = [("Nc", C.c_int32 * 16), ("accepted", C.c_int64 * 16),
                ("err_coll", C.c_int64 * 16), ("skipped_dead", C.c_int64 * 16),
                ("executed", C.c_int64 * 16), ("appended", C.c_int64 * 16)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int32)
        L.lo_ran2.restype = C.c_double
        L.lo_ran2.argtypes = [ip]
        L.lo_seed.restype = C.c_int32
        L.lo_seed.argtypes = [C.c_int32, C.c_int32]
        L.lo_ran2_fill.argtypes = [ip, dp, C.c_int64]
        L.lo_load_particles.restype = C.c_int
        L.lo_load_particles.argtypes = [C.POINTER(Grid), ip, dp, C.c_double, C.c_int32, C.c_int32,
                                        dp, dp, dp, ip, C.POINTER(PSet), dp]
        L.lo_cell_index.restype = C.c_int32
        L.lo_cell_index.argtypes = [C.c_double, C.c_double, C.c_double, C.POINTER(Grid)]
        L.lo_sort_particles_by_cell.argtypes = [C.POINTER(Grid), C.POINTER(PSet), ip, ip]
        L.lo_move_electrostatic.argtypes = [C.POINTER(Grid), C.POINTER(PSet), dp, C.c_double,
                                            C.c_double, C.c_double]
        L.lo_move_boris.argtypes = [C.POINTER(Grid), C.POINTER(PSet), dp, C.c_int, dp, dp,
                                    C.c_double, C.c_double, C.c_double, C.c_int]
        L.lo_move_boris_bgrid.argtypes = [C.POINTER(Grid), C.POINTER(PSet), dp, dp,
                                          C.POINTER(Grid), C.c_double, C.c_double, C.c_double,
                                          C.c_int]
        L.lo_apply_bc.restype = C.c_int
        L.lo_apply_bc.argtypes = [C.POINTER(Grid), C.POINTER(PSet), ip, C.c_double, C.c_double,
                                  C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                  C.c_double, C.c_double, C.c_int32, dp, dp,
                                  C.POINTER(C.c_int64)]
        L.lo_apply_bc_die.restype = C.c_int
        L.lo_apply_bc_die.argtypes = [C.POINTER(Grid), C.POINTER(PSet), ip, C.c_double, C.c_double,
                                      C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                      C.c_double, C.c_double, C.c_int32, C.c_int, ip, dp, dp, dp,
                                      dp, C.POINTER(C.c_int64)]
        L.lo_dielectric_phi.argtypes = [C.POINTER(Grid), ip, ip, C.c_int32, C.c_int32, dp, dp,
                                        C.c_double, dp]
        L.lo_build_kq.argtypes = [C.POINTER(Grid), ip, dp]
        L.lo_deposit.argtypes = [C.POINTER(Grid), C.POINTER(PSet), dp, C.c_double, dp]
        L.lo_deposit_fixed.argtypes = [C.POINTER(Grid), C.POINTER(PSet), dp, C.c_double,
                                       C.c_int32, dp]
        L.lo_fixed_split.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.POINTER(C.c_int64)]
        L.lo_reduce_density.argtypes = [C.POINTER(Grid), ip, C.c_int32, C.c_int32, dp, dp]
        L.lo_build_rho.argtypes = [C.POINTER(Grid), C.c_int32, dp, dp, dp]
        L.lo_pdesolver.restype = C.c_int
        L.lo_pdesolver.argtypes = [C.POINTER(Grid), dp, dp, ip, C.c_int, C.c_int, C.c_double,
                                   C.c_int32, C.c_double, C.c_double, C.c_int32,
                                   C.POINTER(MGStats)]
        L.lo_efield.argtypes = [C.POINTER(Grid), dp, ip, dp]
        L.lo_mg_sor_colour.argtypes = [C.POINTER(Grid), dp, dp, ip, C.c_double, C.c_int,
                                       C.c_double]
        L.lo_mg_periodic.argtypes = [C.POINTER(Grid), dp, ip]
        L.lo_mg_residual.restype = C.c_double
        L.lo_mg_residual.argtypes = [C.POINTER(Grid), dp, dp, ip, C.c_double]
        L.lo_mg_sor_rb.restype = C.c_int
        L.lo_mg_sor_rb.argtypes = [C.POINTER(Grid), dp, dp, ip, C.c_double, C.c_double,
                                   C.c_double, C.c_int, C.c_int, C.c_int, dp]
        L.lo_mg_restrict.argtypes = [C.POINTER(Grid), dp, dp, ip, C.c_double, C.c_int, C.c_int,
                                     dp, ip]
        L.lo_mg_prolong.argtypes = [C.POINTER(Grid), dp, ip, dp]
        L.lo_electron_heating.argtypes = [C.POINTER(Grid), C.POINTER(PSet), C.POINTER(Heating),
                                          ip, dp, dp]
        L.lo_plane_moments.argtypes = [C.POINTER(Grid), C.POINTER(PSet), C.c_double, C.c_double,
                                       C.c_int32, C.c_int32, C.c_int32, dp, dp, dp]
        L.lo_collisions.restype = C.c_int
        L.lo_collisions.argtypes = [C.POINTER(Grid), C.POINTER(PSet), dp, C.POINTER(Chem),
                                    C.c_int32, C.c_double, ip, dp, C.POINTER(MCCStats)]
        _lib = L
    return _lib


def _p(a, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct))


def grid(n, h) -> Grid:
    g = Grid()
    for d in range(3):
        g.n[d] = int(n[d])
        g.h[d] = float(h[d])
    return g


class Particles:
    """Host SoA particle set with a fixed capacity (numpy-owned)."""

    def __init__(self, cap: int):
        self.cap = int(cap)
        self.x = np.zeros(cap); self.y = np.zeros(cap); self.z = np.zeros(cap)
        self.vx = np.zeros(cap); self.vy = np.zeros(cap); self.vz = np.zeros(cap)
        self.cell = np.zeros(cap, np.int32)
        self.dead = np.zeros(cap, np.int8)
        self.s = PSet(_p(self.x), _p(self.y), _p(self.z), _p(self.vx), _p(self.vy), _p(self.vz),
                      _p(self.cell, C.c_int32), _p(self.dead, C.c_int8), 0, self.cap)

    @property
    def n(self) -> int:
        return int(self.s.n)

    @n.setter
    def n(self, v: int):
        self.s.n = int(v)

    def set(self, x, y, z, vx, vy, vz, dead=None):
        m = len(x)
        assert m <= self.cap
        for name, v in zip(("x", "y", "z", "vx", "vy", "vz"), (x, y, z, vx, vy, vz)):
            getattr(self, name)[:m] = v
        self.dead[:] = 0
        if dead is not None:
            self.dead[:m] = dead
        self.s.n = m
        return self

    def arrays(self):
        m = self.n
        return {k: getattr(self, k)[:m].copy() for k in ("x", "y", "z", "vx", "vy", "vz", "cell", "dead")}

    def copy(self) -> "Particles":
        o = Particles(self.cap)
        for k in ("x", "y", "z", "vx", "vy", "vz", "cell", "dead"):
            getattr(o, k)[:] = getattr(self, k)
        o.n = self.n
        return o


def ran2(seed: int, count: int):
    s = C.c_int32(seed)
    out = np.empty(count)
    lib().lo_ran2_fill(C.byref(s), _p(out), count)
    return out, s.value


def seed(iproc: int, rank: int = 0) -> int:
    return lib().lo_seed(iproc, rank)


def load_particles(n, h, bcnd, kq, np_cell, vt0, Nm, ni0, seed_value, cap, x_load=0.0):
    g = grid(n, h)
    nt = len(vt0)
    parts = [Particles(cap) for _ in range(nt)]
    arr = (PSet * nt)(*[p.s for p in parts])
    nodes = bcnd.size
    np_thread = np.zeros(nt * nodes)
    iseed = np.array([seed_value], np.int32)
    st = lib().lo_load_particles(C.byref(g), _p(bcnd, C.c_int32), _p(kq), float(x_load),
                                 int(np_cell), nt, _p(np.asarray(vt0, float)),
                                 _p(np.asarray(Nm, float)), _p(np.asarray(ni0, float)),
                                 _p(iseed, C.c_int32), arr, _p(np_thread))
    if st != 0:
        raise RuntimeError(f"lo_load_particles status {st}")
    for p, s in zip(parts, arr):
        p.s.n = s.n
    return parts, np_thread.reshape(nt, *bcnd.shape), int(iseed[0])


def sort_particles(n, h, p: Particles):
    g = grid(n, h)
    ncells = int(n[0]) * int(n[1]) * int(n[2])
    cc = np.zeros(ncells, np.int32)
    cs = np.zeros(ncells, np.int32)
    lib().lo_sort_particles_by_cell(C.byref(g), C.byref(p.s), _p(cc, C.c_int32), _p(cs, C.c_int32))
    return cc, cs


def move_electrostatic(n, h, p: Particles, E4, q, m, dt):
    g = grid(n, h)
    lib().lo_move_electrostatic(C.byref(g), C.byref(p.s), _p(np.ascontiguousarray(E4)), q, m, dt)


def move_boris(n, h, p: Particles, E4, b_mode, Bconst, B4, q, m, dt, magnetized=True):
    g = grid(n, h)
    Bc = np.asarray(Bconst if Bconst is not None else [0.0, 0.0, 0.0], float)
    B4a = np.ascontiguousarray(B4) if B4 is not None else np.zeros(4)
    lib().lo_move_boris(C.byref(g), C.byref(p.s), _p(np.ascontiguousarray(E4)), int(b_mode),
                        _p(Bc), _p(B4a), q, m, dt, 1 if magnetized else 0)


def move_boris_bgrid(n, h, p, E4, B4, nB, hB, q, m, dt, magnetized=True):
    """Boris push with the B map on its own grid of nB cells of size hB."""
    g, gb = grid(n, h), grid(nB, hB)
    lib().lo_move_boris_bgrid(C.byref(g), C.byref(p.s), _p(E4), _p(np.ascontiguousarray(B4)),
                              C.byref(gb), q, m, dt, int(bool(magnetized)))

def apply_bc(n, h, p: Particles, bcnd, flag_pbc, flag_nmn, ptype, tag_neg, qmacro, mass, Nm, ngrid,
             dtype=None, sum_q_xz=None, sum_q_yz=None):
    """apply_particle_bc_legacy; with dtype (labels 0..ngrid) the dielectric
    branch runs and adds to sum_q_xz (n3+3, n1+3) and sum_q_yz (n3+3, n2+3, 2)."""
    g = grid(n, h)
    p_mac = np.zeros(2 * (ngrid + 1))
    P_loss = np.zeros(1)
    lost = C.c_int64(0)
    xmax, ymax, zmax = (n[d] * h[d] for d in range(3))
    if dtype is None:
        lib().lo_apply_bc(C.byref(g), C.byref(p.s), _p(bcnd, C.c_int32), xmax, ymax, zmax,
                          int(flag_pbc), int(flag_nmn), int(ptype), int(tag_neg), qmacro, mass, Nm,
                          int(ngrid), _p(p_mac), _p(P_loss), C.byref(lost))
    else:
        dt = np.ascontiguousarray(dtype, np.int32)
        assert sum_q_xz.flags.c_contiguous and sum_q_yz.flags.c_contiguous
        lib().lo_apply_bc_die(C.byref(g), C.byref(p.s), _p(bcnd, C.c_int32), xmax, ymax, zmax,
                              int(flag_pbc), int(flag_nmn), int(ptype), int(tag_neg), qmacro, mass,
                              Nm, int(ngrid), 1, _p(dt, C.c_int32), _p(sum_q_xz), _p(sum_q_yz),
                              _p(p_mac), _p(P_loss), C.byref(lost))
    return lost.value, p_mac, float(P_loss[0])


def dielectric_phi(n, h, bcnd, dtype, ngrid, sum_q_xz, sum_q_yz, Ca, phi):
    """apply_dielectric_bc_to_phi on phi (in place); sum_q_* carry ntype first."""
    g = grid(n, h)
    dt = np.ascontiguousarray(dtype, np.int32)
    lib().lo_dielectric_phi(C.byref(g), _p(bcnd, C.c_int32), _p(dt, C.c_int32), int(ngrid),
                            int(sum_q_xz.shape[0]), _p(np.ascontiguousarray(sum_q_xz)),
                            _p(np.ascontiguousarray(sum_q_yz)), float(Ca), _p(phi))


def build_kq(n, bcnd):
    g = grid(n, (1.0, 1.0, 1.0))
    kq = np.zeros(bcnd.shape)
    lib().lo_build_kq(C.byref(g), _p(bcnd, C.c_int32), _p(kq))
    return kq


def deposit(n, h, p: Particles, kq, Nm, np_local=None):
    g = grid(n, h)
    out = np.zeros(kq.shape) if np_local is None else np_local
    assert out.flags.c_contiguous and out.dtype == np.float64
    lib().lo_deposit(C.byref(g), C.byref(p.s), _p(kq), float(Nm), _p(out))
    return out


def deposit_fixed(n, h, p: Particles, kq, Nm, frac_bits, np_local=None):
    g = grid(n, h)
    out = np.zeros(kq.shape) if np_local is None else np_local
    assert out.flags.c_contiguous and out.dtype == np.float64
    lib().lo_deposit_fixed(C.byref(g), C.byref(p.s), _p(kq), float(Nm), int(frac_bits), _p(out))
    return out


def reduce_density(n, bcnd, np_thread):
    """np_thread: (ntype, nodes...) for one virtual thread; stitched in place."""
    g = grid(n, (1.0, 1.0, 1.0))
    nt = np_thread.shape[0]
    np_red = np.zeros_like(np_thread)
    lib().lo_reduce_density(C.byref(g), _p(bcnd, C.c_int32), nt, 1, _p(np_thread), _p(np_red))
    return np_red


def build_rho(n, np_red, charge):
    g = grid(n, (1.0, 1.0, 1.0))
    rho = np.zeros(np_red.shape[1:])
    lib().lo_build_rho(C.byref(g), np_red.shape[0], _p(np_red), _p(np.asarray(charge, float)), _p(rho))
    return rho


def fixed_split(U, ux, uy, uz):
    """The oracle's split of a full weight U over the 8 stencil nodes."""
    a = np.zeros(8, np.int64)
    lib().lo_fixed_split(float(U), float(ux), float(uy), float(uz), _p(a, C.c_int64))
    return a


def pdesolver(n, h, phi, rho, bcnd, flag_pbc=0, flag_nmn=0, eps0=8.854187817e-12, ncycl=20000,
              eps=1e-8, omega=1.4, ng=5):
    g = grid(n, h)
    st = MGStats()
    status = lib().lo_pdesolver(C.byref(g), _p(phi), _p(rho), _p(bcnd, C.c_int32), int(flag_pbc),
                                int(flag_nmn), eps0, int(ncycl), eps, omega, int(ng), C.byref(st))
    return status, {"cycles": st.cycles, "ktot": st.ktot, "res": st.res,
                    "sweeps": [st.sweeps[i] for i in range(int(ng))]}


EPS0 = 8.854187817e-12


def mg_sor_colour(n, h, u, b, bcnd, colour, omega, eps0=EPS0):
    """One colour of the red-black SOR on one level, in place on u."""
    lib().lo_mg_sor_colour(C.byref(grid(n, h)), _p(u), _p(b), _p(bcnd, C.c_int32), eps0,
                           int(colour), float(omega))


def mg_periodic(n, h, u, bcnd):
    lib().lo_mg_periodic(C.byref(grid(n, h)), _p(u), _p(bcnd, C.c_int32))


def mg_residual(n, h, u, b, bcnd, eps0=EPS0):
    """The L1 residual over bcnd == -1 nodes that sor_rb's exit tests use."""
    return lib().lo_mg_residual(C.byref(grid(n, h)), _p(u), _p(b), _p(bcnd, C.c_int32), eps0)


def mg_sor_rb(n, h, u, b, bcnd, omega, eps, kmg, dig, flag_pbc=0, eps0=EPS0):
    """sor_rb on one level, in place on u; returns (sweeps, last residual)."""
    res = C.c_double(0.0)
    k = lib().lo_mg_sor_rb(C.byref(grid(n, h)), _p(u), _p(b), _p(bcnd, C.c_int32), eps0,
                           float(omega), float(eps), int(kmg), int(dig), int(flag_pbc),
                           C.byref(res))
    return k, res.value


def mg_restrict(n, h, u, b, bcnd, flag_pbc=0, flag_nmn=0, eps0=EPS0):
    """One restriction of the fine level's residual; returns the coarse
    level's right-hand side and labels (n/2 cells)."""
    nc = tuple(int(x) // 2 for x in n)
    shape = (nc[2] + 3, nc[1] + 3, nc[0] + 3)
    bc_c = np.zeros(shape, np.int32)
    b_c = np.zeros(shape)
    lib().lo_mg_restrict(C.byref(grid(n, h)), _p(u), _p(b), _p(bcnd, C.c_int32), eps0,
                         int(flag_pbc), int(flag_nmn), _p(b_c), _p(bc_c, C.c_int32))
    return b_c, bc_c


def mg_prolong(n, h, u, bcnd, uc):
    """u += P(uc) on the fine level, in place."""
    lib().lo_mg_prolong(C.byref(grid(n, h)), _p(u), _p(bcnd, C.c_int32), _p(uc))


_omp = None


def pdesolver_threaded(n, h, phi, rho, bcnd, flag_pbc=0, flag_nmn=0, eps0=8.854187817e-12,
                       ncycl=20000, eps=1e-8, omega=1.4, ng=5):
    """pdesolver from liblepic_oracle_omp.so: the same source with its
    multigrid loops on OpenMP threads (OMP_NUM_THREADS, read when the library
    first starts its threads).  Timing only: its residual sums round in
    another order, so it is not the parity checker."""
    global _omp
    if _omp is None:
        build()
        path = os.path.join(_HERE, "_build", "liblepic_oracle_omp.so")
        if not os.path.exists(path):
            build(force=True)
        L = C.CDLL(path)
        dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int32)
        L.lo_pdesolver.restype = C.c_int
        L.lo_pdesolver.argtypes = [C.POINTER(Grid), dp, dp, ip, C.c_int, C.c_int, C.c_double,
                                   C.c_int32, C.c_double, C.c_double, C.c_int32,
                                   C.POINTER(MGStats)]
        _omp = L
    g = grid(n, h)
    st = MGStats()
    status = _omp.lo_pdesolver(C.byref(g), _p(phi), _p(rho), _p(bcnd, C.c_int32), int(flag_pbc),
                               int(flag_nmn), eps0, int(ncycl), eps, omega, int(ng), C.byref(st))
    return status, {"cycles": st.cycles, "ktot": st.ktot, "res": st.res}


def efield(n, h, phi, bcnd):
    g = grid(n, h)
    E4 = np.zeros(phi.shape + (4,))
    lib().lo_efield(C.byref(g), _p(phi), _p(bcnd, C.c_int32), _p(E4))
    return E4


def electron_heating(n, h, p, hp: dict, seed_value: int, P_heat: float = 0.0):
    """Heating of the electron set p (host arrays, updated in place); hp holds
    the lo_heating fields.  Returns (seed, P_heat, (count, energy, vt))."""
    g = grid(n, h)
    k = Heating(**{f: hp[f] for f, _ in Heating._fields_})
    seed = np.array([seed_value], np.int32)
    P = np.array([P_heat])
    out = np.zeros(3)
    lib().lo_electron_heating(C.byref(g), C.byref(p.s), C.byref(k), _p(seed, C.c_int32), _p(P),
                              _p(out))
    return int(seed[0]), float(P[0]), (int(out[0]), float(out[1]), float(out[2]))


def plane_moments(n, h, p, dt, mass, ix_plane, iy_plane, iz_plane, data_xy, data_xz, data_yz):
    """compute_particle_plane_moments_species (mod_planeMoments.f90:13-308);
    data_* are updated in place ((n2+3, n1+3, 5), (n3+3, n1+3, 5), (n3+3, n2+3, 5))."""
    g = grid(n, h)
    for d in (data_xy, data_xz, data_yz):
        assert d.flags.c_contiguous and d.dtype == np.float64
    lib().lo_plane_moments(C.byref(g), C.byref(p.s), float(dt), float(mass), int(ix_plane),
                           int(iy_plane), int(iz_plane), _p(data_xy), _p(data_xz), _p(data_yz))


class ChemTables:
    """Keeps numpy buffers alive behind a Chem struct."""

    def __init__(self, t: dict):
        self.t = {k: (np.ascontiguousarray(v) if isinstance(v, np.ndarray) else v) for k, v in t.items()}
        a = self.t
        self.c = Chem(
            int(a["ntype_tracked"]), int(a["ntype_total"]), _p(a["mass"]), _p(a["vt0"]), _p(a["Nm"]),
            _p(a["neutral_density"]), _p(a["nu_uplim"]), _p(a["p_ncol"], C.c_int32),
            _p(a["sig_list"], C.c_int32), int(a["ncol"]), int(a["npt"]), _p(a["sig_Er"]),
            _p(a["sig"]), _p(a["sig_Eth"]), _p(a["col_info"], C.c_int32), _p(a["sigv_mx"]))


MCC_REC = 27


def collisions_logged(n, h, parts, np_red, chem, ns_coll, dt, seed_value, cap=200000):
    """collisions() with the oracle's event log on: returns its result and
    one row of MCC_REC values per accepted event (see lepic_oracle.h)."""
    L = lib()
    if not hasattr(L, "_mcc_log_typed"):
        L.lo_mcc_set_log.argtypes = [C.POINTER(C.c_double), C.c_int64]
        L.lo_mcc_log_count.restype = C.c_int64
        L._mcc_log_typed = True
    buf = np.zeros((cap, MCC_REC))
    L.lo_mcc_set_log(_p(buf), cap)
    try:
        out = collisions(n, h, parts, np_red, chem, ns_coll, dt, seed_value)
        m = L.lo_mcc_log_count()
    finally:
        L.lo_mcc_set_log(None, 0)
    return out, buf[:m].copy()


def collisions(n, h, parts, np_red, chem: ChemTables, ns_coll, dt, seed_value):
    g = grid(n, h)
    arr = (PSet * len(parts))(*[p.s for p in parts])
    iseed = np.array([seed_value], np.int32)
    Pcoll = np.zeros(chem.t["ntype_total"])
    st = MCCStats()
    status = lib().lo_collisions(C.byref(g), arr, _p(np.ascontiguousarray(np_red)), C.byref(chem.c),
                                 int(ns_coll), float(dt), _p(iseed, C.c_int32), _p(Pcoll), C.byref(st))
    for p, s in zip(parts, arr):
        p.s.n = s.n
    nt = chem.t["ntype_tracked"]
    stats = {"Nc": [st.Nc[i] for i in range(nt)], "accepted": [st.accepted[i] for i in range(nt)],
             "err_coll": [st.err_coll[i] for i in range(nt)],
             "skipped_dead": [st.skipped_dead[i] for i in range(nt)],
             "executed": [st.executed[i] for i in range(nt)],
             "appended": [st.appended[i] for i in range(nt)]}
    return status, int(iseed[0]), Pcoll, stats


class OracleSim:
    """The modular step (mod_simulation.f90:340-472) driven through the oracle
    routines, host arrays, one virtual thread.  Mirrors Simulation."""

    def __init__(self, case, parts=None, np_thread=None, seed_value=123456):
        self.c = case
        self.n, self.h = case.n, case.h
        self.bcnd = np.ascontiguousarray(case.bcnd, np.int32)
        self.kq = build_kq(self.n, self.bcnd)
        self.phi = case.phi0()
        self.rho = np.zeros(self.bcnd.shape)
        tr = case.tracked
        self.ntype = len(tr)
        self.charge = [s.charge for s in tr]
        self.mass = [s.mass for s in tr]
        self.Nm = case.Nm[: self.ntype]
        self.seed = seed_value
        if parts is None:
            ncells = self.n[0] * self.n[1] * self.n[2]
            cap = int(1.3 * abs(case.np_cell) * ncells) + 1024
            ni0 = list(case.ni0) if case.ni0 is not None else [1.0] * self.ntype
            parts, np_thread, self.seed = load_particles(self.n, self.h, self.bcnd, self.kq,
                                                         case.np_cell, case.vt0[: self.ntype],
                                                         self.Nm, ni0, seed_value, cap,
                                                         x_load=case.x_load)
        self.parts, self.np_thread = parts, np.ascontiguousarray(np_thread)
        self.np_red = reduce_density(self.n, self.bcnd, self.np_thread)
        self.p_mac = np.zeros((self.ntype, 2 * (case.ngrid + 1)))
        self.P_loss = np.zeros(self.ntype)
        self.chem = None
        if case.chemistry is not None:
            t = dict(case.chemistry)
            t["Nm"] = np.asarray(case.Nm, float)
            self.chem = ChemTables(t)
            self.Pcoll = np.zeros(t["ntype_total"])
        self.tag_neg = -1
        for i, q in enumerate(self.charge):
            if q < 0.0 and i != 0:
                self.tag_neg = i + 1
                break
        self.last_mg = None
        self.lost = [0] * self.ntype
        self.last_mcc = None
        self.P_heat = 0.0
        self.R_ahp_seen = 0.0
        if case.flag_die:
            n1, n2, n3 = (int(v) for v in self.n)
            self.sum_q_xz = np.zeros((self.ntype, n3 + 3, n1 + 3))
            self.sum_q_yz = np.zeros((self.ntype, n3 + 3, n2 + 3, 2))
        self.last_heating = None

    def advance_one_step(self, istep):
        c, n, h = self.c, self.n, self.h
        if istep % c.nb_step_sort == 1:
            for p in self.parts:
                if p.n > 1:
                    sort_particles(n, h, p)
        self.np_red = reduce_density(n, self.bcnd, self.np_thread)
        rho = build_rho(n, self.np_red, self.charge)
        self.rho = rho
        if c.flag_die:
            dielectric_phi(n, h, self.bcnd, c.dtype, c.ngrid, self.sum_q_xz, self.sum_q_yz,
                           c.Ca * h[0] * h[2], self.phi)
        st, self.last_mg = pdesolver(n, h, self.phi, rho, self.bcnd, c.flag_pbc, c.flag_nmn,
                                     ncycl=c.ncycl, eps=c.eps, omega=c.omega, ng=c.ng)
        if st != 0:
            raise RuntimeError(f"pdesolver status {st}")
        E4 = efield(n, h, self.phi, self.bcnd)
        for s, p in enumerate(self.parts):
            q, m = self.charge[s], self.mass[s]
            if c.B_map is not None and c.B_grid is not None:
                move_boris_bgrid(n, h, p, E4, c.B_map, c.B_grid[0], c.B_grid[1], q, m, c.dt,
                                 q <= 0 or c.ions_magnetized)
            elif c.B_map is not None:
                move_boris(n, h, p, E4, 2, None, c.B_map, q, m, c.dt, q <= 0 or c.ions_magnetized)
            elif c.B_uniform is not None:
                move_boris(n, h, p, E4, 1, c.B_uniform, None, q, m, c.dt, q <= 0 or c.ions_magnetized)
            else:
                move_electrostatic(n, h, p, E4, q, m, c.dt)
        if c.flag_die:
            self.sum_q_xz[:] = 0.0
            self.sum_q_yz[:] = 0.0
        for s, p in enumerate(self.parts):
            die = (c.dtype, self.sum_q_xz[s], self.sum_q_yz[s]) if c.flag_die else (None, None, None)
            lost, pm, Pl = apply_bc(n, h, p, self.bcnd, c.flag_pbc, c.flag_nmn, s + 1, self.tag_neg,
                                    self.Nm[s] * self.charge[s], self.mass[s], self.Nm[s], c.ngrid,
                                    *die)
            self.lost[s] = lost
            self.p_mac[s] += pm
            self.P_loss[s] += Pl
        if self.chem is not None and c.nb_step_collisions > 0 and istep % c.nb_step_collisions == 0:
            st, self.seed, Pc, self.last_mcc = collisions(n, h, self.parts, self.np_red, self.chem,
                                                          c.nb_step_collisions, c.dt, self.seed)
            if st != 0:
                raise RuntimeError(f"collisions status {st}")
            self.Pcoll += Pc
        hc = c.heating
        if hc is not None and hc["nb_step_heating"] > 0 and istep % hc["nb_step_heating"] == 0:
            R_app = hc["R_ahp"] if hc["flag_circxh"] == 1 else self.R_ahp_seen
            hp = dict(nudt=hc["nudt"], Pabs=hc["Pabs"], nu_h=hc["nu_h"], xl_pow=hc["xl_pow"],
                      xr_pow=hc["xr_pow"], flag_circxh=hc["flag_circxh"],
                      flag_ahp=hc["flag_ahp"], R_ahp_moments=self.R_ahp_seen, R_ahp=R_app,
                      ymax=hc["ymax"], zmax=hc["zmax"], Nm=self.Nm[0], mass=self.mass[0],
                      vt0=c.vt0[0])
            self.seed, self.P_heat, self.last_heating = electron_heating(
                n, h, self.parts[0], hp, self.seed, self.P_heat)
            self.R_ahp_seen = R_app
        self.np_thread = np.zeros_like(self.np_thread)
        for s, p in enumerate(self.parts):
            deposit(n, h, p, self.kq, self.Nm[s], self.np_thread[s])
        self.np_red = reduce_density(n, self.bcnd, self.np_thread)


# ---------------------------------------------------------------------------
# The FP32 particle phase (cell id + FP32 in-cell offsets + FP32 velocities)
# ---------------------------------------------------------------------------
class PSet32(C.Structure):
    _fields_ = [("cell", C.POINTER(C.c_int32)),
                ("fx", C.POINTER(C.c_float)), ("fy", C.POINTER(C.c_float)),
                ("fz", C.POINTER(C.c_float)), ("vx", C.POINTER(C.c_float)),
                ("vy", C.POINTER(C.c_float)), ("vz", C.POINTER(C.c_float)),
                ("n", C.c_int64), ("cap", C.c_int64)]


FIELDS32 = ("cell", "fx", "fy", "fz", "vx", "vy", "vz")


class Particles32:
    """Host FP32 particle set (numpy-owned): cell (int32, -1 = no particle),
    offsets fx fy fz in cell widths and velocities, float32."""

    def __init__(self, cap: int):
        self.cap = int(cap)
        self.cell = np.full(cap, -1, np.int32)
        for k in FIELDS32[1:]:
            setattr(self, k, np.zeros(cap, np.float32))
        self.s = PSet32(_p(self.cell, C.c_int32), *[_p(getattr(self, k), C.c_float)
                                                   for k in FIELDS32[1:]], 0, self.cap)

    @property
    def n(self) -> int:
        return int(self.s.n)

    @n.setter
    def n(self, v: int):
        self.s.n = int(v)

    def set(self, cell, fx, fy, fz, vx, vy, vz):
        m = len(cell)
        assert m <= self.cap
        for k, v in zip(FIELDS32, (cell, fx, fy, fz, vx, vy, vz)):
            getattr(self, k)[:m] = v
        self.s.n = m
        return self

    def arrays(self):
        m = self.n
        return {k: getattr(self, k)[:m].copy() for k in FIELDS32}

    def copy(self) -> "Particles32":
        o = Particles32(self.cap)
        for k in FIELDS32:
            getattr(o, k)[:] = getattr(self, k)
        o.n = self.n
        return o


_lib32 = None


def lib32():
    global _lib32
    if _lib32 is None:
        L = lib()
        dp, ip, lp = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int64)
        fp = C.POINTER(C.c_float)
        L.lo32_round4.argtypes = [dp, C.c_int64, fp]
        L.lo32_from64.restype = C.c_int64
        L.lo32_from64.argtypes = [C.POINTER(Grid), C.POINTER(PSet), C.POINTER(PSet32)]
        L.lo32_to64.argtypes = [C.POINTER(Grid), C.POINTER(PSet32), C.POINTER(PSet)]
        L.lo32_push.argtypes = [C.POINTER(Grid), C.POINTER(PSet32), fp, C.c_int, dp, fp, C.c_int,
                                C.c_double, C.c_double, C.c_double, ip, C.c_int, C.c_int, C.c_int,
                                C.c_int, C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32,
                                lp, dp, dp, lp, lp]
        L.lo32_deposit.argtypes = [C.POINTER(Grid), C.POINTER(PSet32), C.c_double, C.c_int32, lp]
        L.lo32_fixed_density.argtypes = [C.POINTER(Grid), lp, dp, C.c_int32, dp]
        L.lo32_sort.argtypes = [C.POINTER(Grid), C.POINTER(PSet32), C.POINTER(PSet32)]
        _lib32 = L
    return _lib32


def round4(F4):
    """A 4-per-node field rounded to float32."""
    a = np.ascontiguousarray(F4, np.float64)
    out = np.zeros(a.shape, np.float32)
    lib32().lo32_round4(_p(a), a.size, _p(out, C.c_float))
    return out


def from64(n, h, p: Particles, cap=None) -> tuple:
    """FP64 set -> FP32 set (slot for slot); returns (set, outside count)."""
    q = Particles32(cap or p.cap)
    out = lib32().lo32_from64(C.byref(grid(n, h)), C.byref(p.s), C.byref(q.s))
    return q, int(out)


def to64(n, h, q: Particles32, cap=None) -> Particles:
    p = Particles(cap or q.cap)
    lib32().lo32_to64(C.byref(grid(n, h)), C.byref(q.s), C.byref(p.s))
    return p


def push32(n, h, p: Particles32, E4f, bmode, Bconst, B4f, q, m, dt, bcnd, flag_pbc, flag_nmn,
           ptype, tag_neg, qmacro, mass, Nm, ngrid, frac_bits=0, dep=None, magnetized=True):
    """The FP32 particle phase in place; dep (int64 per node) receives the
    fixed-point deposit when given.  Returns (lost, escaped, p_mac, P_loss)."""
    g = grid(n, h)
    Bc = np.asarray(Bconst if Bconst is not None else [0.0, 0.0, 0.0], float)
    B4a = np.ascontiguousarray(B4f, np.float32) if B4f is not None else np.zeros(4, np.float32)
    E = np.ascontiguousarray(E4f, np.float32)
    p_mac = np.zeros(2 * (ngrid + 1))
    P_loss = np.zeros(1)
    lost, esc = C.c_int64(0), C.c_int64(0)
    if dep is not None:
        assert dep.dtype == np.int64 and dep.flags.c_contiguous
    lib32().lo32_push(C.byref(g), C.byref(p.s), _p(E, C.c_float), int(bmode), _p(Bc),
                      _p(B4a, C.c_float), 1 if magnetized else 0, q, m, dt,
                      _p(np.ascontiguousarray(bcnd, np.int32), C.c_int32), int(flag_pbc),
                      int(flag_nmn), int(ptype), int(tag_neg), qmacro, mass, Nm, int(ngrid),
                      int(frac_bits), _p(dep, C.c_int64) if dep is not None else None,
                      _p(p_mac), _p(P_loss), C.byref(lost), C.byref(esc))
    return lost.value, esc.value, p_mac, float(P_loss[0])


def deposit32(n, h, p: Particles32, Nm, frac_bits, dep=None):
    g = grid(n, h)
    out = np.zeros(tuple(int(d) + 3 for d in n[::-1]), np.int64) if dep is None else dep
    lib32().lo32_deposit(C.byref(g), C.byref(p.s), float(Nm), int(frac_bits), _p(out, C.c_int64))
    return out


def fixed_density(n, dep, kq, frac_bits, np_local=None):
    g = grid(n, (1.0, 1.0, 1.0))
    out = np.zeros(kq.shape) if np_local is None else np_local
    lib32().lo32_fixed_density(C.byref(g), _p(np.ascontiguousarray(dep), C.c_int64), _p(kq),
                               int(frac_bits), _p(out))
    return out


def sort32(n, h, p: Particles32) -> Particles32:
    q = Particles32(p.cap)
    lib32().lo32_sort(C.byref(grid(n, h)), C.byref(p.s), C.byref(q.s))
    return q
