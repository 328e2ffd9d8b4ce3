"""Run decks in the reference's format: conditions.inp, geometry.inp,
boundary.inp and the chemistry file they name.

read_conditions          read_input_cfg      modules/config/mod_readConditions.f90:73-224
read_geometry            read_geometry_file  modules/geometry/mod_boundary.f90:19-40
read_boundary_potentials                     modules/geometry/mod_boundary.f90:43-54
generate_boundary        generate_boundary   modules/geometry/mod_generateBoundary.f90:74-433
build_bfield             MagneticField%build_from_cfg and its profiles
                                             modules/fields/mod_magneticField.f90:70-196, 336-567
sim_params               SimParams%build     modules/simulation/mod_simParams.f90:56-169
load_deck / Deck.case    the set-up order of State (modules/core/mod_state.f90:197-246)

All of this is one-time host set-up.  The files are read with Fortran
list-directed semantics (records, blank/comma/slash separators, r*c repeats,
'!' is not a comment: a line's trailing text is skipped only because the read
statement is already satisfied).  Lengths are in cm in the files and in m
after set-up, as in the reference.  Grid arrays are numpy arrays of shape
(n3+3, n2+3, n1+3): node (i,j,k) of the 0:n+2 index space is element [k, j, i].
"""
from __future__ import annotations

import dataclasses
import math
import os
import re
import warnings

import numpy as np

from . import config as cfg

NBMAX = 5        # mod_config.f90:7
NM_RG = 5        # mod_config.f90:8
MAX_SEGMENTS = 100   # mod_boundary.f90:31-35
MODULAR_STEPS = 150005   # modules/main/main.f90:14


class DeckError(ValueError):
    """A deck the reference's reader or set-up stops on."""


# ---------------------------------------------------------------------------
# Fortran list-directed input
# ---------------------------------------------------------------------------

_NULL = object()
_SLASH = object()
_INT = re.compile(r"[+-]?\d+$")
_REAL = re.compile(r"([+-]?(?:\d+\.?\d*|\.\d+))(?:[eEdDqQ]([+-]?\d+)|([+-]\d+))?$")


def _record_values(line: str) -> list:
    """The values of one record: strings, _NULL for null values (two commas in
    a row), _SLASH where a slash ends the statement."""
    out, i, n = [], 0, len(line)
    after_value = False       # a value (or null) was just read; a comma closes it
    while i < n:
        c = line[i]
        if c in " \t":
            i += 1
            continue
        if c == ",":
            if not after_value:
                out.append(_NULL)
            after_value = False
            i += 1
            continue
        if c == "/":
            out.append(_SLASH)
            return out
        if c in "'\"":
            j, s = i + 1, []
            while j < n:
                if line[j] == c:
                    if j + 1 < n and line[j + 1] == c:
                        s.append(c)
                        j += 2
                        continue
                    break
                s.append(line[j])
                j += 1
            out.append("".join(s))
            i = j + 1
            after_value = True
            continue
        j = i
        while j < n and line[j] not in " \t,/":
            j += 1
        tok = line[i:j]
        m = re.match(r"(\d+)\*(.*)$", tok)
        if m:
            r, rest = int(m.group(1)), m.group(2)
            out.extend([_NULL if rest == "" else rest] * r)
        else:
            out.append(tok)
        i = j
        after_value = True
    return out


class _Statement:
    """One list-directed READ: values are taken from the current record and
    the following ones as needed; the rest of the last record is skipped."""

    def __init__(self, rec: "_Records"):
        self.rec = rec
        self.vals = []
        self.done = False           # a slash ended the statement

    def _next(self):
        if self.done:
            return _NULL
        while not self.vals:
            self.vals = self.rec._take()     # raises EOFError at the end of the file
        v = self.vals.pop(0)
        if v is _SLASH:
            self.done = True
            return _NULL
        return v

    def int(self, default=0):
        v = self._next()
        if v is _NULL:
            return default
        if not _INT.match(v):
            raise DeckError(f"bad integer {v!r} (record {self.rec.i})")
        return int(v)

    def real(self, default=0.0):
        v = self._next()
        if v is _NULL:
            return default
        m = _REAL.match(v)
        if not m:
            raise DeckError(f"bad real {v!r} (record {self.rec.i})")
        exp = m.group(2) if m.group(2) is not None else m.group(3)
        return float(m.group(1) + ("e" + exp if exp is not None else ""))

    def char(self, length: int, default=""):
        v = self._next()
        if v is _NULL:
            return default
        return v[:length]


class _Records:
    def __init__(self, text: str):
        self.lines = text.splitlines()
        self.i = 0

    def _take(self) -> list:
        if self.i >= len(self.lines):
            raise EOFError
        line = self.lines[self.i]
        self.i += 1
        return _record_values(line)

    def statement(self) -> _Statement:
        return _Statement(self)


def _text(path_or_text) -> str:
    """A deck's text: a string holding a newline is the text itself, anything
    else names a file that must exist."""
    if isinstance(path_or_text, str) and "\n" in path_or_text:
        return path_or_text
    if not os.path.isfile(path_or_text):
        raise DeckError(f"no such deck file: {path_or_text}")
    with open(path_or_text) as f:
        return f.read()


# ---------------------------------------------------------------------------
# conditions.inp
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class BSource:
    """One magnetic-field record of conditions.inp (B0 in T, lengths in m)."""
    file: str = " "
    name: str = ""
    scale: float = 0.0
    B0: float = 0.0
    info: str = ""
    dL: float = 0.0
    x0: float = 0.0
    y0: float = 0.0
    z0: float = 0.0


@dataclasses.dataclass
class Conditions:
    """The Config fields read_input_cfg fills (mod_config.f90:11-101)."""
    rname: str = ""
    Te: float = 0.0
    ng: int = 0
    eps: float = 0.0
    omega: float = 0.0
    n: tuple = (0, 0, 0)
    xl_pow: float = 0.0
    xr_pow: float = 0.0
    yl_pow: float = 0.0
    yr_pow: float = 0.0
    zl_pow: float = 0.0
    zr_pow: float = 0.0
    x_load: float = 0.0
    flag_ahp: int = 0
    nB: int = 0
    flag_B_pos: int = 0
    n_B: tuple = (1, 1, 1)
    flag_B: int = 0
    B: list = dataclasses.field(default_factory=lambda: [BSource() for _ in range(NBMAX)])
    Pabs: float = 0.0
    nu_h: float = 0.0
    flag_heat: int = 0
    I_inj: float = 0.0
    opt_inj: int = 0
    flag_inj: int = 0
    tmax: float = 0.0
    kt: float = 0.0
    n0: float = 0.0
    ngas: float = 0.0
    np_cell: int = 0
    ngrid: int = 0
    flag_avg3D: int = 1
    nsav: int = 0
    nbak: int = 0
    cnt_plt: int = 0
    tseq_init: float = 0.0
    tseq_final: float = 0.0
    tseq: float = 0.0
    k_eps0: float = 0.0
    flag_convP: int = 0
    flag_pdf: int = 0
    ptype_pdf: int = 0
    yl_rg: float = 0.0
    yr_rg: float = 0.0
    zl_rg: float = 0.0
    zr_rg: float = 0.0
    n_rg: int = 0
    xl_rg: list = dataclasses.field(default_factory=lambda: [0.0] * NM_RG)
    xr_rg: list = dataclasses.field(default_factory=lambda: [0.0] * NM_RG)
    flag_restart: int = 0
    omp_rank_max: int = 1
    mpi_rank_max: int = 1
    np_dup: float = 1.0
    flag_grd: int = 0
    Lhy: float = 0.0
    Lhz: float = 0.0
    nhy: int = 0
    nhz: int = 0
    ind_g: int = 0
    jne: float = 0.0
    THm: float = 0.0
    num_grd: int = 0
    Ca: float = 0.0
    gam_sec: float = 0.0
    igrid_sec: int = 0
    flag_RFpot: int = 0
    phi0_RF: float = 0.0
    f0_RF: float = 0.0
    phi1_RF: float = 0.0
    f1_RF: float = 0.0


def read_conditions(path_or_text) -> Conditions:
    """read_input_cfg (mod_readConditions.f90:73-224): one READ per record,
    in the file's order, then the same flags and unit conversions.  A record
    that does not hold what its READ asks for raises DeckError, as the
    reference's err= branch stops the run."""
    rec = _Records(_text(path_or_text))
    c = Conditions()
    try:
        c.rname = rec.statement().char(20).rstrip()
        c.Te = rec.statement().real()
        c.ng = rec.statement().int()
        s = rec.statement()
        c.eps, c.omega = s.real(), s.real()
        s = rec.statement()
        c.n = (s.int(), s.int(), s.int())
        s = rec.statement()
        pw = [s.real() * 1.0e-2 for _ in range(6)]
        c.xl_pow, c.xr_pow, c.yl_pow, c.yr_pow, c.zl_pow, c.zr_pow = pw
        c.x_load = rec.statement().real() * 1.0e-2
        if c.yr_pow < 0.0:
            c.flag_ahp = 1
            c.yr_pow = abs(c.yr_pow)
        c.nB = rec.statement().int()
        c.flag_B_pos = 1 if c.nB < 0 else 0
        c.nB = abs(c.nB)
        s = rec.statement()
        n_B = [s.int(), s.int(), s.int()]
        if c.nB > NBMAX:
            raise DeckError(f"{c.nB} magnetic-field records, at most {NBMAX}")
        for iB in range(c.nB):
            s = rec.statement()
            b = c.B[iB]
            b.file, b.name, b.scale, b.B0 = s.char(1), s.char(20).rstrip(), s.real(), s.real()
            b.info, b.dL, b.x0, b.y0, b.z0 = s.char(2), s.real(), s.real(), s.real(), s.real()
            b.B0 = b.B0 * 1.0e-4
            if b.file not in ("s", "S"):
                c.flag_B = 1
        if n_B[0] <= 1 or n_B[1] <= 1 or n_B[2] <= 1 or c.flag_B == 0 or c.nB == 0:
            n_B = [1, 1, 1]
        c.n_B = tuple(n_B)
        if c.flag_B == 0:
            c.nB = 0
        for b in c.B:
            b.dL, b.x0, b.y0, b.z0 = b.dL * 1.0e-2, b.x0 * 1.0e-2, b.y0 * 1.0e-2, b.z0 * 1.0e-2
        s = rec.statement()
        c.Pabs, c.nu_h = s.real(), s.real()
        c.flag_heat = 0 if c.nu_h < 0.0 else 1
        s = rec.statement()
        c.I_inj, c.opt_inj = s.real(), s.int()
        c.flag_inj = 0 if c.I_inj < 0.0 else 1
        c.tmax = rec.statement().real()
        c.kt = rec.statement().real()
        c.n0 = rec.statement().real()
        c.ngas = rec.statement().real()
        c.np_cell = rec.statement().int()
        c.ngrid = rec.statement().int()
        c.flag_avg3D = 1
        if c.ngrid < 0:
            c.flag_avg3D = 0
            c.ngrid = abs(c.ngrid)
        c.nsav = rec.statement().int()
        s = rec.statement()
        c.nbak, c.cnt_plt, c.tseq_init, c.tseq_final = s.int(), s.int(), s.real(), s.real()
        c.tseq = abs(c.tseq_final - c.tseq_init) / float(c.cnt_plt) if c.cnt_plt != 0 else math.inf
        c.k_eps0 = rec.statement().real()
        if c.k_eps0 < 0.0:
            c.flag_convP = 1
            c.k_eps0 = abs(c.k_eps0)
        s = rec.statement()
        ans = s.char(1)
        c.ptype_pdf = s.int()
        c.yl_rg, c.yr_rg, c.zl_rg, c.zr_rg = s.real(), s.real(), s.real(), s.real()
        c.n_rg = s.int()
        if c.n_rg > NM_RG:
            raise DeckError(f"{c.n_rg} pdf regions, at most {NM_RG} (nm_rg)")
        for i in range(c.n_rg):
            c.xl_rg[i], c.xr_rg[i] = s.real(), s.real()
        c.zl_rg, c.zr_rg = c.zl_rg * 1.0e-2, c.zr_rg * 1.0e-2
        c.yl_rg, c.yr_rg = c.yl_rg * 1.0e-2, c.yr_rg * 1.0e-2
        c.xl_rg = [v * 1.0e-2 for v in c.xl_rg]
        c.xr_rg = [v * 1.0e-2 for v in c.xr_rg]
        c.flag_pdf = 1 if ans in ("y", "Y") else 0
        s = rec.statement()
        ans = s.char(1)
        c.omp_rank_max, c.mpi_rank_max, c.np_dup = s.int(), s.int(), s.real()
        c.flag_restart = {"y": 1, "Y": 1, "e": 2, "E": 2}.get(ans, 0)
        if c.flag_restart == 0:
            c.np_dup = 1.0
        s = rec.statement()
        ans = s.char(1)
        c.Lhy, c.Lhz, c.nhy, c.nhz, c.ind_g = s.real(), s.real(), s.int(), s.int(), s.int()
        c.flag_grd = 1 if ans in ("y", "Y") else 0
        s = rec.statement()
        c.jne, c.THm, c.num_grd, c.Ca = s.real(), s.real(), s.int(), s.real()
        s = rec.statement()
        c.gam_sec, c.igrid_sec = s.real(), s.int()
        ans = s.char(1)
        c.phi0_RF, c.f0_RF, c.phi1_RF, c.f1_RF = s.real(), s.real(), s.real(), s.real()
        c.flag_RFpot = 1 if ans in ("y", "Y") else 0
        if c.I_inj > 0.0 and c.gam_sec > 0.0:
            raise DeckError("gam_sec > 0 is incompatible with I_inj > 0")
        while True:
            end = rec.statement().char(3)
            if end in ("END", "end"):
                break
    except EOFError:
        raise DeckError("conditions file ends before its END record") from None
    return c


# ---------------------------------------------------------------------------
# geometry.inp and boundary.inp
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class Geometry:
    """GeometrySegments (mod_generateBoundary.f90:63-68), lengths in cm."""
    xmax: float
    ymax: float
    zmax: float
    segments: list     # (xl, yl, zl, xr, yr, zr, ind)


def read_geometry(path_or_text) -> Geometry:
    """read_geometry_file (mod_boundary.f90:19-40): the box, then up to 100
    segments of seven values; the end of the file ends the list."""
    rec = _Records(_text(path_or_text))
    try:
        s = rec.statement()
        box = (s.real(), s.real(), s.real())
    except EOFError:
        raise DeckError("geometry file holds no box size") from None
    segs = []
    for _ in range(MAX_SEGMENTS):
        try:
            s = rec.statement()
            seg = (s.real(), s.real(), s.real(), s.real(), s.real(), s.real(), s.int())
        except EOFError:
            break
        segs.append(seg)
    return Geometry(box[0], box[1], box[2], segs)


def read_boundary_potentials(path_or_text, ngrid: int) -> list:
    """read_boundary_potentials (mod_boundary.f90:43-54): V(1:ngrid), one
    record each, 0 past the end of the file."""
    rec = _Records(_text(path_or_text))
    V = [0.0] * ngrid
    for ig in range(ngrid):
        try:
            V[ig] = rec.statement().real()
        except EOFError:
            break
    return V


# ---------------------------------------------------------------------------
# generate_boundary
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class Boundary:
    """The arrays and BoundaryOutputs of generate_boundary (lengths in m)."""
    bcnd: np.ndarray
    u: np.ndarray
    h: tuple
    dtype: np.ndarray          # 0:ngrid
    xmax: float = 0.0
    ymax: float = 0.0
    zmax: float = 0.0
    Lgy: float = 0.0
    Lgz: float = 0.0
    Sg: float = 0.0
    xg1: float = 0.0
    ixg: int = 0
    flag_circxh: int = 0
    flag_pbc: int = 0
    flag_pbcz: int = 0
    flag_nmn: int = 0
    flag_die: int = 0
    ig_die: list = dataclasses.field(default_factory=lambda: [0] * 5)
    zg_sec: list = dataclasses.field(default_factory=lambda: [0.0, 0.0])
    n_cath: int = 0
    dir_sec: int = 1
    every: int = 1


def _nint(x: float) -> int:
    """Fortran NINT: nearest integer, halves away from zero."""
    return int(math.copysign(math.floor(abs(x) + 0.5), x))


def generate_boundary(n, V, ngrid: int, geo: Geometry, xl_pow: float, xr_pow: float,
                      flag_grd: int = 0, ind_g: int = 0, nhy: int = 0, nhz: int = 0,
                      Lhy: float = 0.0, Lhz: float = 0.0, gam_sec: float = 0.0,
                      opt_inj: int = 0, igrid_sec: int = 0) -> Boundary:
    """generate_boundary (mod_generateBoundary.f90:74-433) without the .mco
    plane files: segments painted in file order, then the Neumann wall, the
    periodic fix-up along z, |dtype| and the grid holes.  V holds the
    electrode potentials V(1:ngrid)."""
    n1, n2, n3 = (int(v) for v in n)
    bcnd = np.zeros((n3 + 3, n2 + 3, n1 + 3), np.int32)
    u = np.zeros(bcnd.shape)
    dtype = np.zeros(ngrid + 1, np.int64)
    out = Boundary(bcnd, u, (0.0, 0.0, 0.0), dtype)
    mx = max(n1, n2, n3)
    out.every = 8 if mx > 2048 else 4 if mx > 1024 else 2 if mx > 512 else 1
    xmax, ymax, zmax = geo.xmax, geo.ymax, geo.zmax
    if xmax < 0.0:
        out.flag_nmn = 1
    xmax = abs(xmax)
    flag_circx_tmp = 1 if ymax < 0.0 else 0
    ymax = abs(ymax)
    flag_circz = 1 if zmax < 0.0 else 0
    zmax = abs(zmax)
    h1, h2, h3 = xmax / n1, ymax / n2, zmax / n3
    R = 0.0
    cnt_yz_planes = 0
    igl = igr = None

    def coords(lo, hi, h):
        return (np.arange(lo, hi + 1) - 1).astype(np.float64) * h

    for (xl, yl, zl, xr, yr, zr, ind) in geo.segments:
        a = abs(ind)
        if a > ngrid:
            raise DeckError("insufficient number of wall labels in the boundary file")
        dtype[a] = 1
        flag_circx = flag_circx_tmp
        if yl < 0.0:
            flag_circx = 0 if flag_circx_tmp == 1 else 1
        yl = abs(yl)
        if flag_circx == 1:
            R = min((yr - yl) / 2.0, (zr - zl) / 2.0)
            if xl <= xl_pow and xr >= xr_pow:
                out.flag_circxh = 1
        if flag_circz == 1:
            R = min((xr - xl) / 2.0, (yr - yl) / 2.0)
        if ind == 0:
            out.flag_pbc = 1
        if xr < 0.0:
            xr = abs(xr)
            dtype[a] = 3 if cnt_yz_planes == 0 else 4
            cnt_yz_planes += 1
            out.flag_die = 1
            out.ig_die[dtype[a] - 1] = int(xl / h1) + 1
        if yr < 0.0:
            yr = abs(yr)
            dtype[a] = 2
            out.flag_die = 1
            out.ig_die[1] = int(yl / h2) + 1
        if zr < 0.0:
            zr = abs(zr)
            dtype[a] = -dtype[a]
            out.flag_pbc = 1
            out.flag_pbcz = 1
        ixl = int(xl / h1) + 1
        ixl = 0 if ixl <= 1 else ixl
        ixr = int(xr / h1) + 1
        ixr = n1 + 2 if ixr >= n1 else ixr
        iyl = int(yl / h2) + 1
        iyl = 0 if iyl <= 1 else iyl
        iyr = int(yr / h2) + 1
        iyr = n2 + 2 if iyr >= n2 else iyr
        izl = int(zl / h3) + 1
        izl = 0 if izl <= 1 else izl
        izr = int(zr / h3) + 1
        izr = n3 + 2 if izr >= n3 else izr
        out.dir_sec = 1
        if (gam_sec > 0.0 or abs(opt_inj) == 4) and igrid_sec == a:
            if zr != zl:
                raise DeckError("secondary emission is modelled along z only")
            if zl == 0.0:
                out.zg_sec[0] = zr
                out.n_cath += 1
            if zr == zmax:
                out.zg_sec[1] = zl
                out.dir_sec = -1
                out.n_cath += 1
        if a == ind_g:
            igl, igr = ixl, ixr
            out.Lgy = min((iyr - iyl) * h2, ymax)
            out.Lgz = min((izr - izl) * h3, zmax)
            out.ixg = _nint(float(igl + igr) / 2.0)
            out.xg1 = (ixl - 1) * h1
        if ind == 0:
            ixl, ixr = max(ixl, 2), min(ixr, n1)
        if ind >= 0:
            if ixl <= ixr:
                bcnd[:, :, ixl:ixr + 1] = ind
                if ind > 0:
                    u[:, :, ixl:ixr + 1] = V[ind - 1]
            ixl, ixr = max(ixl, 2), min(ixr, n1)
            iyl, iyr = max(iyl, 2), min(iyr, n2)
            izl, izr = max(izl, 2), min(izr, n3)
        if ixl > ixr or iyl > iyr or izl > izr:
            continue
        x = coords(ixl, ixr, h1)[None, None, :]
        y = coords(iyl, iyr, h2)[None, :, None]
        z = coords(izl, izr, h3)[:, None, None]
        inside = np.ones((izr - izl + 1, iyr - iyl + 1, ixr - ixl + 1), bool)
        if flag_circz == 1 and (ind > 0 or ind < 0):
            dx, dy = x - xmax / 2.0, y - ymax / 2.0
            inside = inside & ~((dx * dx + dy * dy) > R * R)
        if flag_circx == 1 and (ind > 0 or ind < 0):
            dy, dz = y - ymax / 2.0, z - zmax / 2.0
            inside = inside & ~((dy * dy + dz * dz) > R * R)
        blk = (slice(izl, izr + 1), slice(iyl, iyr + 1), slice(ixl, ixr + 1))
        if ind >= 0:
            bcnd[blk] = np.where(inside, -1, bcnd[blk])
            u[blk] = np.where(inside, 0.0, u[blk])
        else:
            bcnd[blk] = np.where(inside, a, bcnd[blk])
            if a > 0:
                u[blk] = np.where(inside, V[a - 1], u[blk])

    if out.flag_nmn == 1:
        bcnd[:, :, 0:2] = -2
        u[:, :, 0:2] = 0.0
    if out.flag_pbc == 1:
        # ig = bcnd(ix,iy,1) is a label there, or -2 on a Neumann wall where
        # bcnd(ix,iy,2) is -2 as well, so the test below never holds for it
        ig = bcnd[1]
        lab = np.where(ig >= 0, ig, 0)
        per = (ig >= 0) & (dtype[lab] < 0) & (bcnd[2] == -1)
        for kk in (0, 1, n3 + 1, n3 + 2):
            bcnd[kk][per] = 0
    dtype[:] = np.abs(dtype)

    Sh = 0.0
    if flag_grd == 1:
        if igl is None:
            raise DeckError(f"no segment carries the grid label {ind_g} the holes are drawn on")
        iyh = _nint(Lhy / 2.0 / h2) + 1
        izh = _nint(Lhz / 2.0 / h3) + 1
        if Lhy > ymax:
            raise DeckError("Lhy > ymax")
        if Lhz > zmax:
            raise DeckError("Lhz > zmax")
        if Lhy == Lhz:
            R = Lhy / 2.0
            Sh = math.pi * R ** 2
        else:
            Sh = Lhy * Lhz
        for cnt_hz in range(1, nhz + 1):
            zd = -(nhz - 1) * 3.0 * Lhz / 4.0 + (cnt_hz - 1) * 3.0 * Lhz / 2.0 + zmax / 2.0
            izl = _nint(zd / h3) + 1 - izh
            izr = izl + 2 * izh
            izl = max(izl, 1)
            izr = min(izr, n3 + 1)
            for cnt_hy in range(1, nhy + 1):
                yd = -(nhy - 1) * 3.0 * Lhy / 4.0 + (cnt_hy - 1) * 3.0 * Lhy / 2.0 + ymax / 2.0
                iyl = _nint(yd / h2) + 1 - iyh
                iyr = iyl + 2 * iyh
                iyl = max(iyl, 1)
                iyr = min(iyr, n2 + 1)
                if iyl > iyr or izl > izr or igl > igr:
                    continue
                iy = np.arange(iyl, iyr + 1)[None, :]
                iz = np.arange(izl, izr + 1)[:, None]
                y = (iy - 1).astype(np.float64) * h2
                z = (iz - 1).astype(np.float64) * h3
                val = np.where((iy == 1) | (iy == n2 + 1) | (iz == 1) | (iz == n3 + 1), 0, -1)
                hit = np.ones(val.shape, bool)
                if Lhy == Lhz:
                    dy, dz = y - yd, z - zd
                    hit = (dy * dy + dz * dz) <= R * R
                blk = (slice(izl, izr + 1), slice(iyl, iyr + 1), slice(igl, igr + 1))
                bcnd[blk] = np.where(hit[:, :, None], val[:, :, None], bcnd[blk])
                u[blk] = np.where(hit[:, :, None], 0.0, u[blk])

    out.h = (h1 * 1.0e-2, h2 * 1.0e-2, h3 * 1.0e-2)
    out.xmax, out.ymax, out.zmax = xmax * 1.0e-2, ymax * 1.0e-2, zmax * 1.0e-2
    out.xg1 *= 1.0e-2
    out.Lgy *= 1.0e-2
    out.Lgz *= 1.0e-2
    if flag_grd == 1:
        out.Sg = out.Lgy * out.Lgz - (nhy - 1) * (nhz - 1) * Sh * 1.0e-4
    else:
        out.Sg = out.Lgy * out.Lgz
    out.zg_sec = [v * 1.0e-2 for v in out.zg_sec]
    return out


# ---------------------------------------------------------------------------
# Magnetic field (mod_magneticField.f90)
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class MagneticMap:
    """Bi of MagneticField: shape (nzB+3, nyB+3, nxB+3, 4), components Bx, By,
    Bz and |B| in T; h_B in m."""
    n_B: tuple
    h_B: tuple
    Bi: np.ndarray
    Bmax: float = 0.0
    xBm: tuple = (0, 0, 0)
    flag_thr: int = 0


def _b_dir(info: str) -> int:
    """find_B_dir (mod_magneticField.f90:336-346), 0-based component."""
    if "y" in info or "Y" in info:
        return 1
    if "z" in info or "Z" in info:
        return 2
    return 0


def _node_coord(count: int, h: float) -> np.ndarray:
    return (np.arange(count) - 1).astype(np.float64) * h


def _gaussian(Bi, h_B, B0, x0, sigma, comp):
    """gaussian_Bfield (mod_magneticField.f90:403-423); exp per node column."""
    if sigma < 0.0:
        return
    x = _node_coord(Bi.shape[2], h_B[0])
    prof = np.array([B0 * math.exp(-((xi - x0) ** 2) / (2.0 * sigma ** 2)) for xi in x])
    Bi[:, :, :, comp] = Bi[:, :, :, comp] + prof[None, None, :]


def _pg_current(Bi, n_B, h_B, B0, comp):
    """PG_current (mod_magneticField.f90:426-447)."""
    L = max(h_B[0] * float(n_B[0]), 1.0e-12)
    x = _node_coord(Bi.shape[2], h_B[0])
    f = 1.0 / (1.0 + (x / L) ** 2)
    Bi[:, :, :, comp] = Bi[:, :, :, comp] + (B0 * f)[None, None, :]


def _ee_magnets(Bi, h_B, B0, x0, z0, dL):
    """EE_magnets (mod_magneticField.f90:450-476); exp per (x, z) node."""
    if dL <= 0.0:
        return
    x = _node_coord(Bi.shape[2], h_B[0])
    z = _node_coord(Bi.shape[0], h_B[2])
    ax = (x - x0) / dL
    az = (z - z0) / dL
    g = np.array([[math.exp(-(axi * axi + azk * azk)) for axi in ax] for azk in az])
    bx = B0 * (-2.0 * ax[None, :] / dL) * g
    bz = B0 * (+2.0 * az[:, None] / dL) * g
    Bi[:, :, :, 0] = Bi[:, :, :, 0] + bx[:, None, :]
    Bi[:, :, :, 2] = Bi[:, :, :, 2] + bz[:, None, :]


def _bx_magnet(x, y, a):
    return 0.5 * (a * x / (x ** 2 + (y + a) ** 2) - a * x / (x ** 2 + (y - a) ** 2))


def _by_magnet(x, y, a):
    return 0.5 * (a * (y + a) / (x ** 2 + (y + a) ** 2) - a * (y - a) / (x ** 2 + (y - a) ** 2))


def _hall_thruster(Bi, h_B, B0, x0, a):
    """Bfield_Hall_thruster (mod_magneticField.f90:479-547)."""
    d = 1.5e-2
    y0 = 2.0e-2
    alpha = 0.75
    B1 = abs(B0) / (1.0 - alpha * a ** 2 / (4.0 * d ** 2 + a ** 2))
    B2 = alpha * B1
    if B0 < 0.0:
        y0 = 0.0
    x = _node_coord(Bi.shape[2], h_B[0])[None, :]
    y = _node_coord(Bi.shape[1], h_B[1])[:, None]
    bx = B1 * _bx_magnet(x - x0, y - y0, a) - B2 * _bx_magnet(x - x0 + 2.0 * d, y - y0, a)
    by = B1 * _by_magnet(x - x0, y - y0, a) - B2 * _by_magnet(x - x0 + 2.0 * d, y - y0, a)
    Bi[:, :, :, 0] = Bi[:, :, :, 0] + bx[None, :, :]
    Bi[:, :, :, 1] = Bi[:, :, :, 1] + by[None, :, :]
    if B0 < 0.0:
        Bi[:, :, :, 0] = 0.0
        Bi[:, :, :, 1] = Bi[1, 1, :, 1][None, None, :]


def _read_map(Bi, path: str, comp: int, scale: float):
    """read_Bfield_map (mod_magneticField.f90:349-400): an optional header
    record of two or three integers, then one value per record, x fastest."""
    with open(path) as f:
        lines = f.read().splitlines()
    start = 0
    if lines:
        head = _record_values(lines[0])
        if len(head) >= 2 and all(isinstance(v, str) and _INT.match(v) for v in head[:2]):
            start = 1
    rec = _Records("\n".join(lines[start:]))
    nz, ny, nx = Bi.shape[:3]
    vals = np.empty(nz * ny * nx)
    try:
        for t in range(vals.size):
            vals[t] = rec.statement().real()
    except EOFError:
        raise DeckError(f"B map {path} ends early") from None
    Bi[:, :, :, comp] = Bi[:, :, :, comp] + scale * vals.reshape(nz, ny, nx)


def build_bfield(c: Conditions, xmax: float, ymax: float, zmax: float,
                 directory: str = ".") -> MagneticMap:
    """MagneticField%build_from_cfg (mod_magneticField.f90:103-165) on the
    domain extents in m (as generate_boundary leaves them).  'y' records read
    <name>Bx.dat, <name>By.dat and <name>Bz.dat from directory."""
    nB = [int(v) for v in c.n_B]
    if nB[0] <= 1 or nB[1] <= 1 or nB[2] <= 1:
        nB = [1, 1, 1]
    Bi = np.zeros((nB[2] + 3, nB[1] + 3, nB[0] + 3, 4))
    h_B = (xmax / float(nB[0]), ymax / float(nB[1]), zmax / float(nB[2]))
    m = MagneticMap(tuple(nB), h_B, Bi)
    if c.flag_B != 0 and c.nB > 0:
        for b in c.B[:min(c.nB, NBMAX)]:
            if b.file in ("s", "S"):
                continue
            if b.file in ("y", "Y"):
                for comp, ax in enumerate("xyz"):
                    _read_map(Bi, os.path.join(directory, f"{b.name}B{ax}.dat"), comp, b.scale)
            elif b.file in ("n", "N"):
                comp = _b_dir(b.info)
                key = b.name[:1]
                if key in ("g", "G"):
                    _gaussian(Bi, h_B, b.B0, b.x0, b.dL, comp)
                elif key in ("p", "P"):
                    _pg_current(Bi, nB, h_B, b.B0, comp)
                elif key in ("e", "E"):
                    _ee_magnets(Bi, h_B, b.B0, b.x0, b.z0, b.dL)
                elif key in ("t", "T"):
                    _hall_thruster(Bi, h_B, b.B0, b.x0, b.dL)
                    m.flag_thr = 1
    # compute_norm (mod_magneticField.f90:169-190): |B| and the first node of
    # the maximum in x-fastest order
    Bi[..., 3] = np.sqrt(Bi[..., 0] ** 2 + Bi[..., 1] ** 2 + Bi[..., 2] ** 2)
    flat = Bi[..., 3].reshape(-1)
    if flat.size and flat.max() > 0.0:
        t = int(np.argmax(flat))
        k, r = divmod(t, Bi.shape[1] * Bi.shape[2])
        j, i = divmod(r, Bi.shape[2])
        m.Bmax, m.xBm = float(flat[t]), (i, j, k)
    return m


# ---------------------------------------------------------------------------
# Derived step parameters (mod_simParams.f90)
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class SimParams:
    vt0: list
    Nm: list
    dt: float
    lbd_d: float
    wp: float
    nb_step_sort: int = 10
    nb_step_collisions: int = 10
    nb_step_heating: int = 4
    nb_step_averaging: int = 0
    nseq: int = 0
    tseq: float = 0.0
    nudt: float = 0.0
    nu_uplim: list = dataclasses.field(default_factory=list)


def sim_params(c: Conditions, h, species, ntype: int) -> SimParams:
    """SimParams%build (mod_simParams.f90:56-169) for the species list of the
    chemistry (the electron first, its temperature from conditions.inp);
    ntype is the number of charged species."""
    e = species[0]
    if c.n0 <= 0.0:
        raise DeckError("n0 must be > 0")
    if abs(e.charge) <= 0.0 or e.mass <= 0.0:
        raise DeckError("invalid charge or mass of species 1")
    lbd_d = math.sqrt((cfg.EPS0 * c.Te) / (c.n0 * abs(e.charge)))
    wp = math.sqrt(c.n0 * e.charge ** 2 / (cfg.EPS0 * e.mass))
    if c.np_cell == 0:
        raise DeckError("np_cell must be nonzero")
    vol = h[0] * h[1] * h[2]
    vt0, Nm = [], []
    for k, s in enumerate(species):
        T = c.Te if k == 0 else s.T_eV
        mabs = abs(s.mass)
        vt0.append(math.sqrt(2.0 * cfg.QE * T / mabs) if mabs > 0.0 and T > 0.0 else 0.0)
        dens = c.n0 if abs(s.charge) > 0.0 else c.ngas
        Nm.append(dens * h[0] * h[1] * h[2] / abs(float(c.np_cell)))
    hmin = min(h[0], min(h[1], h[2]))
    if vt0[0] <= 0.0:
        raise DeckError("vt0(1) <= 0, cannot compute dt")
    dt = c.kt * hmin / vt0[0]
    p = SimParams(vt0, Nm, dt, lbd_d, wp)
    p.nb_step_sort = 10
    p.nb_step_collisions = 1 * p.nb_step_sort
    p.nb_step_heating = 4
    if 0.05 <= c.kt < 0.1:
        p.nb_step_heating = 20
    if c.kt < 0.05:
        p.nb_step_heating = 40
    p.nb_step_averaging = 5 * p.nb_step_sort if c.nbak != 0 else c.nsav
    p.tseq = c.tseq
    if p.tseq > 0.0 and math.isfinite(p.tseq):
        p.nseq = max(_nint((p.tseq / dt) / float(c.nsav)), 1) * c.nsav
        p.tseq = float(p.nseq) * dt
    p.nudt = c.nu_h * (float(p.nb_step_heating) * dt)
    p.nu_uplim = [cfg.NU_UPLIM_ELECTRON] + [cfg.NU_UPLIM_OTHER] * max(ntype - 1, 0)
    del vol
    return p


# ---------------------------------------------------------------------------
# A whole deck
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class Deck:
    directory: str
    conditions: Conditions
    geometry: Geometry
    V: list
    boundary: Boundary
    chemistry: object          # chemistry.Chemistry
    bfield: MagneticMap
    params: SimParams

    def legacy_only(self) -> list:
        """Features this deck sets that only the legacy driver (Src/main.f90)
        runs: the modular step (mod_simulation.f90:340-472) has no particle
        injection, secondary emission, RF or iterative electrode potentials,
        negative-ion surface production or restart, so a deck asking for them
        runs without them there, and here."""
        c, out = self.conditions, []
        if c.flag_inj:
            out.append("particle injection")
        if c.gam_sec > 0.0:
            out.append("secondary emission")
        if c.flag_RFpot:
            out.append("RF electrode potential")
        if c.flag_convP:
            out.append("iterative electrode potential")
        if c.flag_restart:
            out.append("restart")
        if c.jne > 0.0:
            out.append("negative-ion surface production")
        return out

    def case(self, steps: int = MODULAR_STEPS, strict: bool = False) -> cfg.Case:
        """The deck as a Case for Simulation, run as the modular driver runs
        it.  Features only the legacy driver has (legacy_only) are left off
        with a warning, or refused with DeckError when strict."""
        miss = self.legacy_only()
        if miss:
            if strict:
                raise DeckError("deck asks for " + ", ".join(miss) +
                                " (legacy driver only)")
            warnings.warn("left off, as in the modular driver: " + ", ".join(miss), stacklevel=2)
        c, b, chem = self.conditions, self.boundary, self.chemistry
        from . import chemistry as chm
        species = chem.species
        tracked = [s for s in species if s.tracked]
        B_uniform = B_map = B_grid = None
        if c.flag_B:
            # part_expmover.f90:143-175: n_B(1) = 1 is a constant field, a map
            # of the field grid's size shares its stencil, any other map has
            # its own (main.f90:368-370)
            if self.bfield.n_B[0] == 1:
                B_uniform = tuple(float(v) for v in self.bfield.Bi[1, 1, 1, :3])
            else:
                B_map = np.ascontiguousarray(self.bfield.Bi)
                if tuple(self.bfield.n_B) != tuple(c.n):
                    B_grid = (tuple(self.bfield.n_B), tuple(self.bfield.h_B))
        case = cfg.Case(
            name=os.path.basename(os.path.normpath(self.directory)) or "deck",
            n=tuple(c.n), h=b.h, bcnd=b.bcnd, ngrid=c.ngrid,
            electrode_V={k + 1: v for k, v in enumerate(self.V)},
            species=species, n0=c.n0, np_cell=c.np_cell, dt=self.params.dt, steps=steps,
            eps=c.eps, omega=c.omega, ng=c.ng, flag_pbc=b.flag_pbc, flag_nmn=b.flag_nmn,
            B_uniform=B_uniform, B_map=B_map, B_grid=B_grid, ions_magnetized=(c.flag_B_pos == 0),
            ngas=c.ngas, nb_step_sort=self.params.nb_step_sort,
            nb_step_collisions=self.params.nb_step_collisions)
        case.phi_init = b.u
        case.ni0 = [chem.density[k] for k, s in enumerate(species) if s.tracked][:len(tracked)]
        case.x_load = c.x_load
        case.flag_die, case.dtype, case.Ca = b.flag_die, np.array(b.dtype, np.int32), c.Ca
        # the modular driver heats every nb_step_heating steps whatever the
        # deck says (mod_simulation.f90:441-447); nu_h < 0 gives nudt < 0,
        # so nothing is re-drawn but the window's draws are still taken.
        # The heating reads Config's flag_circxh, which keeps its default 0:
        # build_boundary copies generate_boundary's other outputs to the
        # domain but not this one (mod_boundary.f90:114-131).
        case.heating = dict(
            nb_step_heating=self.params.nb_step_heating, nudt=self.params.nudt, Pabs=c.Pabs,
            nu_h=c.nu_h, xl_pow=c.xl_pow, xr_pow=c.xr_pow, flag_circxh=0,
            flag_ahp=c.flag_ahp, R_ahp=c.yr_pow - b.ymax / 2.0, ymax=b.ymax, zmax=b.zmax)
        if chem.reactions and c.ngas > 0.0:
            case.chemistry = chm.build_tables(chem, c.ngas)
        return case


def load_deck(directory: str, conditions: str = "conditions.inp",
              geometry: str = "geometry.inp", boundary: str = "boundary.inp",
              chemistry_dir: str | None = None) -> Deck:
    """Reads a deck directory in the reference's set-up order
    (mod_state.f90:197-246): conditions, chemistry (the file conditions.inp
    names, in chemistry_dir, by default the deck's directory), boundary
    potentials, geometry, generate_boundary, magnetic field, step parameters."""
    from . import chemistry as chm
    c = read_conditions(os.path.join(directory, conditions))
    rpath = os.path.join(chemistry_dir or directory, c.rname)
    if not os.path.isfile(rpath):
        raise DeckError(f"chemistry file {c.rname} not found in {chemistry_dir or directory}")
    chem = chm.read_reactions(rpath, c.ngas, Te=c.Te)
    V = read_boundary_potentials(os.path.join(directory, boundary), c.ngrid)
    geo = read_geometry(os.path.join(directory, geometry))
    b = generate_boundary(c.n, V, c.ngrid, geo, c.xl_pow, c.xr_pow, flag_grd=c.flag_grd,
                          ind_g=c.ind_g, nhy=c.nhy, nhz=c.nhz, Lhy=c.Lhy, Lhz=c.Lhz,
                          gam_sec=c.gam_sec, opt_inj=c.opt_inj, igrid_sec=c.igrid_sec)
    bf = build_bfield(c, b.xmax, b.ymax, b.zmax, directory)
    ntype = sum(1 for s in chem.species if s.tracked)
    p = sim_params(c, b.h, chem.species, ntype)
    return Deck(directory, c, geo, V, b, chem, bf, p)


def write_conditions(c: Conditions) -> str:
    """A conditions.inp holding c, record for record in the order
    read_conditions reads (numbers with 17 significant digits; lengths back
    in cm, B0 back in G), so that a deck written here reads back to c."""
    def f(v):
        return f"{float(v):.17g}"
    cm = lambda v: f(v / 1.0e-2)   # noqa: E731
    rec = [c.rname or "none.dat", f(c.Te), str(c.ng), f"{f(c.eps)} {f(c.omega)}",
           " ".join(str(int(v)) for v in c.n),
           " ".join(cm(v) for v in (c.xl_pow, c.xr_pow, c.yl_pow,
                                    -c.yr_pow if c.flag_ahp else c.yr_pow, c.zl_pow, c.zr_pow)),
           cm(c.x_load), str(-c.nB if c.flag_B_pos else c.nB),
           " ".join(str(int(v)) for v in c.n_B)]
    for b in c.B[:c.nB]:
        rec.append(f"{b.file} {b.name} {f(b.scale)} {f(b.B0 / 1.0e-4)} {b.info} {cm(b.dL)} "
                   f"{cm(b.x0)} {cm(b.y0)} {cm(b.z0)}")
    rec += [f"{f(c.Pabs)} {f(c.nu_h)}", f"{f(c.I_inj)} {c.opt_inj}", f(c.tmax), f(c.kt),
            f(c.n0), f(c.ngas), str(c.np_cell), str(c.ngrid if c.flag_avg3D else -c.ngrid),
            str(c.nsav), f"{c.nbak} {c.cnt_plt} {f(c.tseq_init)} {f(c.tseq_final)} Z",
            f(-c.k_eps0 if c.flag_convP else c.k_eps0)]
    rg = " ".join(f"{cm(c.xl_rg[i])} {cm(c.xr_rg[i])}" for i in range(c.n_rg))
    rec.append(f"{'yes' if c.flag_pdf else 'no'} {c.ptype_pdf} {cm(c.yl_rg)} {cm(c.yr_rg)} "
               f"{cm(c.zl_rg)} {cm(c.zr_rg)} {c.n_rg} {rg}".rstrip())
    rec.append(f"{ {0: 'no', 1: 'yes', 2: 'external'}[c.flag_restart]} {c.omp_rank_max} "
               f"{c.mpi_rank_max} {f(c.np_dup)}")
    rec.append(f"{'yes' if c.flag_grd else 'no'} {f(c.Lhy)} {f(c.Lhz)} {c.nhy} {c.nhz} {c.ind_g}")
    rec.append(f"{f(c.jne)} {f(c.THm)} {c.num_grd} {f(c.Ca)}")
    rec.append(f"{f(c.gam_sec)} {c.igrid_sec} {'yes' if c.flag_RFpot else 'no'} {f(c.phi0_RF)} "
               f"{f(c.f0_RF)} {f(c.phi1_RF)} {f(c.f1_RF)}")
    rec.append("END")
    return "\n".join(rec) + "\n"


def write_geometry(geo: Geometry) -> str:
    def f(v):
        return f"{float(v):.17g}"
    out = [f"{f(geo.xmax)} {f(geo.ymax)} {f(geo.zmax)}"]
    out += [" ".join(f(v) for v in s[:6]) + f" {int(s[6])}" for s in geo.segments]
    return "\n".join(out) + "\n"


def write_boundary_potentials(V) -> str:
    return "\n".join(f"{float(v):.17g}" for v in V) + "\n"
