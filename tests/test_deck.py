"""Run-deck front end (deck.py): the list-directed reader, the reference's
own decks (input_dir/*), boundary generation against what each geometry
fixes, and the magnetic-field profiles against the reference's
MagneticField built from its own sources (oracle/_ref).

generate_boundary's module needs MPI (mod_utils.f90:7-17 calls MPI_BARRIER),
so it is not built here; its checks are closed forms: label layouts derived
by hand from the segment rules for the reference's geometries, and circles
and holes checked against exact rational distances."""
import ctypes as C
import dataclasses
import os
from fractions import Fraction

import numpy as np
import pytest

from essaitemporaire_lepic_module_amd import chemistry as CH
from essaitemporaire_lepic_module_amd import config as cfg
from essaitemporaire_lepic_module_amd import deck as D

REF = "/root/reference/input_dir"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_ref", "libref_lepic.so")
need_ref = pytest.mark.skipif(not os.path.isdir(REF), reason="reference decks absent")
need_lib = pytest.mark.skipif(not os.path.exists(LIB), reason="oracle/_ref not built")

DECKS_OK = ["1D_limit", "3D_zoom", "ITER", "Metris", "Penning", "Thruster", "example_sans"]
DECKS_BAD = {"3D_box": 23, "Dpace_TRIUMF": 28}   # record the reference's reader stops on


# ---------------------------------------------------------------------------
# list-directed reads
# ---------------------------------------------------------------------------

def _st(text):
    return D._Records(text).statement()


def test_list_directed_values():
    s = _st("1, 2 ,3\n")
    assert [s.int(), s.int(), s.int()] == [1, 2, 3]
    s = _st("1,,3\n")
    assert [s.int(), s.int(7), s.int()] == [1, 7, 3]
    s = _st("1 2 / 9\n")
    assert [s.int(), s.int(), s.int(-1)] == [1, 2, -1]
    s = _st("3*2.5 1\n")
    assert [s.real() for _ in range(4)] == [2.5, 2.5, 2.5, 1.0]
    s = _st("'a b' c\n")
    assert s.char(20) == "a b" and s.char(1) == "c"
    reals = _st("1.d-7 8.D10 30.e-6 -.5 1.5+3 7 2.\n")
    assert [reals.real() for _ in range(7)] == [1e-7, 8e10, 30e-6, -0.5, 1.5e3, 7.0, 2.0]
    with pytest.raises(D.DeckError):
        _st("1.\n").int()
    with pytest.raises(D.DeckError):
        _st("! a comment\n").int()


def test_list_directed_records():
    """A READ takes further records when it needs them and skips the rest of
    its last record; blank records hold no values."""
    rec = D._Records("1 2\n\n3 4\n5 ! five\n")
    s = rec.statement()
    assert [s.int(), s.int(), s.int()] == [1, 2, 3]
    assert rec.statement().int() == 5
    with pytest.raises(EOFError):
        rec.statement().int()


# ---------------------------------------------------------------------------
# the reference's decks
# ---------------------------------------------------------------------------

def _tokens(path):
    with open(path) as f:
        return [ln.replace(",", " ").split() for ln in f.read().splitlines()]


def _num(t):
    return float(t.replace("d", "e").replace("D", "e"))


@need_ref
@pytest.mark.parametrize("name", DECKS_OK)
def test_reference_conditions_records(name):
    """Fields land on the records the reference reads them from (record
    numbers follow the number of B-field records)."""
    path = os.path.join(REF, name, "conditions.inp")
    c = D.read_conditions(path)
    t = _tokens(path)
    assert c.rname == t[0][0]
    assert c.Te == _num(t[1][0]) and c.ng == int(t[2][0])
    assert (c.eps, c.omega) == (_num(t[3][0]), _num(t[3][1]))
    assert c.n == tuple(int(v) for v in t[4][:3])
    nb = abs(int(t[7][0]))
    r = 9 + nb                                  # 0-based record of Pabs
    assert c.Pabs == _num(t[r][0])
    assert c.I_inj == _num(t[r + 1][0]) and c.opt_inj == int(t[r + 1][1])
    assert c.tmax == _num(t[r + 2][0]) and c.kt == _num(t[r + 3][0])
    assert c.n0 == _num(t[r + 4][0]) and c.ngas == _num(t[r + 5][0])
    assert c.np_cell == int(t[r + 6][0])
    assert c.ngrid == abs(int(t[r + 7][0])) and c.flag_avg3D == (0 if int(t[r + 7][0]) < 0 else 1)
    assert c.nsav == int(t[r + 8][0])
    assert c.flag_heat == (0 if c.nu_h < 0 else 1)
    assert c.flag_inj == (0 if c.I_inj < 0 else 1)
    assert c.flag_grd == (1 if t[r + 13][0][0] in "yY" else 0)
    # lengths converted to m
    assert c.x_load == pytest.approx(_num(t[6][0]) * 1e-2, rel=1e-15)
    assert c.xl_pow == pytest.approx(_num(t[5][0]) * 1e-2, rel=1e-15)


@need_ref
@pytest.mark.parametrize("name,record", sorted(DECKS_BAD.items()))
def test_reference_conditions_rejected(name, record):
    """Older decks whose records no longer hold what the reader asks for
    (3D_box: '!' read as the holes' grid index; Dpace_TRIUMF: 'Secondary'
    read as phi0_RF) stop, as the reference's err= branch does."""
    with pytest.raises(D.DeckError, match=f"record {record}"):
        D.read_conditions(os.path.join(REF, name, "conditions.inp"))


@need_ref
@pytest.mark.parametrize("name", DECKS_OK + sorted(DECKS_BAD))
def test_reference_geometry_and_potentials(name):
    gpath = os.path.join(REF, name, "geometry.inp")
    geo = D.read_geometry(gpath)
    t = [r for r in _tokens(gpath) if r]
    assert (geo.xmax, geo.ymax, geo.zmax) == tuple(_num(v) for v in t[0][:3])
    assert len(geo.segments) == len(t) - 1
    for seg, r in zip(geo.segments, t[1:]):
        assert seg[:6] == tuple(_num(v) for v in r[:6]) and seg[6] == int(r[6])
    bpath = os.path.join(REF, name, "boundary.inp")
    bt = [r for r in _tokens(bpath) if r]
    V = D.read_boundary_potentials(bpath, len(bt) + 2)
    assert V == [_num(r[0]) for r in bt] + [0.0, 0.0]


# ---------------------------------------------------------------------------
# generate_boundary
# ---------------------------------------------------------------------------

def _geo(text):
    return D.read_geometry(text)


BOX_GEO = "10. 10. 10.\n0. 0. 0. 10. 10. 10. 2\n10. 0. 0. 10.0 10. 10. -3\n0. 0. 0. 0. 10. 10. -1\n"
SLAB_GEO = "10. 2.5 2.5\n0. 0. 0. 10. 2.5 2.5 0\n10. 0. 0. 10. 2.5 2.5 -2\n0. 0. 0. 0. 2.5 2.5 -1\n"
PENNING_GEO = ("10. 10. -53.3\n0. 0. 0. 10. 10. 53.3 1\n0. 0. 53.3 10. 10. 53.3 -2\n"
               "0. 0. 0. 10. 10. 0. -3\n")


@pytest.mark.parametrize("n", [(8, 8, 8), (64, 64, 64), (12, 10, 6)])
def test_boundary_box(n):
    """3D_box's geometry: the whole box painted label 2 with plasma on nodes
    2..n of every axis, then the x = L plane (label 3) and the x = 0 plane
    (label 1) over their ghost layers."""
    V = [0.5, 5.0, -3.0, 0.0]
    b = D.generate_boundary(n, V, 4, _geo(BOX_GEO), 0.0, 0.1)
    n1, n2, n3 = n
    want = np.full(cfg.node_shape(n), 2, np.int32)
    want[2:n3 + 1, 2:n2 + 1, 2:n1 + 1] = -1
    want[:, :, 0:2] = 1
    want[:, :, n1 + 1:] = 3
    np.testing.assert_array_equal(b.bcnd, want)
    u = np.where(want > 0, np.array([0.0] + V)[np.maximum(want, 0)], 0.0)
    np.testing.assert_array_equal(b.u, u)
    assert (b.flag_pbc, b.flag_nmn, b.flag_die) == (0, 0, 0)
    assert list(b.dtype) == [0, 1, 1, 1, 0]
    assert b.h == (0.1 / n1, 0.1 / n2, 0.1 / n3)


@pytest.mark.parametrize("n", [(16, 8, 8), (128, 32, 32)])
def test_boundary_periodic_slab(n):
    """1D_limit's geometry: label 0 (periodic) segment over x nodes 2..n with
    plasma inside, electrodes 1 and 2 on the x ends."""
    b = D.generate_boundary(n, [0.0, 7.0], 2, _geo(SLAB_GEO), 0.0, 0.1)
    n1, n2, n3 = n
    want = np.zeros(cfg.node_shape(n), np.int32)
    want[2:n3 + 1, 2:n2 + 1, 2:n1 + 1] = -1
    want[:, :, 0:2] = 1
    want[:, :, n1 + 1:] = 2
    np.testing.assert_array_equal(b.bcnd, want)
    assert (b.flag_pbc, b.flag_pbcz) == (1, 0)
    assert list(b.dtype) == [1, 1, 1]
    np.testing.assert_array_equal(b.u[:, :, n1 + 1:], 7.0)


def _inside(i, h, c, R):
    """Exact test of (i-1)h - c against R for the doubles given: -1 inside
    by a clear margin, 1 outside, 0 too close to call."""
    d2 = sum((Fraction(ii - 1) * Fraction(hh) - Fraction(cc)) ** 2 for ii, hh, cc in zip(i, h, c))
    r2 = Fraction(R) ** 2
    if abs(d2 - r2) <= r2 * Fraction(1, 10 ** 9):
        return 0
    return -1 if d2 < r2 else 1


def test_boundary_cylinder_along_z():
    """Penning's geometry (zmax < 0: circular cross-section in x-y): plasma
    is the interior nodes within R = 5 cm of the axis; the end plates carry
    their labels inside the circle only."""
    n = (24, 24, 40)
    b = D.generate_boundary(n, [0.0, -35.0, -35.0], 3, _geo(PENNING_GEO), 0.0, 0.1)
    n1, n2, n3 = n
    h = (10.0 / n1, 10.0 / n2)
    for j in range(n2 + 3):
        for i in range(n1 + 3):
            s = _inside((i, j), h, (5.0, 5.0), 5.0)
            if s == 0:
                continue
            col = b.bcnd[:, j, i]
            interior = 2 <= i <= n1 and 2 <= j <= n2
            if s < 0:
                assert (col[[0, 1]] == 3).all() and (col[[n3 + 1, n3 + 2]] == 2).all()
                assert (col[2:n3 + 1] == (-1 if interior else 1)).all(), (i, j)
            else:
                assert (col == 1).all(), (i, j)


def test_boundary_holes():
    """3D_zoom's holes: a circle of radius Lh/2 = 0.4 cm at the centre of the
    y-z plane, cut through the x nodes of the grid segment (label 2)."""
    geo = _geo("-3.2 1.2 1.2\n0. 0. 0. 3.2 1.2 1.2 1\n0. 0. 0. 3.2 1.2 1.2 0\n"
               "2.65 0. 0. 2.85 1.2 1.2 -2\n")
    n = (256, 96, 96)
    b = D.generate_boundary(n, [400.0, 0.0, 0.0, 0.0, 0.0], 5, geo, 0.0, 0.002, flag_grd=1,
                            ind_g=2, nhy=1, nhz=1, Lhy=0.8, Lhz=0.8)
    n1, n2, n3 = n
    h1, h2, h3 = 3.2 / n1, 1.2 / n2, 1.2 / n3
    igl, igr = int(2.65 / h1) + 1, int(2.85 / h1) + 1
    assert (igl, igr) == (212, 229)        # 2.65 / 0.0125 rounds to 211.99999999999997
    for k in range(1, n3 + 2):
        for j in range(1, n2 + 2):
            s = _inside((j, k), (h2, h3), (0.6, 0.6), 0.4)
            if s == 0:
                continue
            row = b.bcnd[k, j, igl:igr + 1]
            if s < 0:
                assert (row == -1).all(), (j, k)
            else:
                assert (row == 2).all(), (j, k)
    assert b.flag_nmn == 1 and (b.bcnd[:, :, 0:2] == -2).all()
    assert b.Sg == pytest.approx(0.012 * 0.012, rel=1e-15)
    assert b.ixg == 221


@need_ref
def test_boundary_thruster_dielectrics():
    """Thruster's segments set by hand from the rules: yr < 0 -> dtype 2,
    xr < 0 -> dtype 3 (first y-z dielectric), zr < 0 -> periodic along z."""
    geo = D.read_geometry(os.path.join(REF, "Thruster", "geometry.inp"))
    b = D.generate_boundary((256, 512, 192), [0.0, 300.0, 0.0, 0.0, 0.0], 5, geo, 0.0015, 0.0075)
    assert list(b.dtype) == [0, 1, 1, 2, 3, 2]
    assert (b.flag_die, b.flag_pbc, b.flag_pbcz) == (1, 1, 1)
    assert b.ig_die == [0, 1, 95, 0, 0]
    # periodic along z: no dielectric label left on the z ghost planes above
    # a plasma node
    plasma2 = b.bcnd[2] == -1
    for k in (0, 1, 193, 194):
        assert not np.isin(b.bcnd[k][plasma2], [3, 4, 5]).any()


@need_ref
@pytest.mark.parametrize("name", DECKS_OK)
def test_reference_deck_boundaries(name):
    """Every label the boundary holds is a label of the deck; electrode nodes
    carry their potential, plasma nodes 0; kq is 2 off the plasma."""
    d = D.load_deck(os.path.join(REF, name), chemistry_dir=_chem_dir(name))
    b, V = d.boundary, d.V
    labels = np.unique(b.bcnd)
    assert labels.min() >= -2 and labels.max() <= d.conditions.ngrid
    assert (b.u[b.bcnd == -1] == 0.0).all()
    for lab in labels[labels >= 1]:
        assert (b.u[b.bcnd == lab] == V[lab - 1]).all(), lab
    np.testing.assert_array_equal(cfg.build_kq(b.bcnd), np.where(b.bcnd != -1, 2.0, 1.0))


def _chem_dir(name):
    c = D.read_conditions(os.path.join(REF, name, "conditions.inp"))
    return None if os.path.isfile(os.path.join(REF, name, c.rname)) else REF


# ---------------------------------------------------------------------------
# magnetic field against the reference's MagneticField
# ---------------------------------------------------------------------------

def _ref_bfield(c, box):
    lib = C.CDLL(LIB)
    nb = [int(v) for v in c.n_B]
    if min(nb) <= 1:
        nb = [1, 1, 1]
    out = np.zeros((nb[2] + 3, nb[1] + 3, nb[0] + 3, 4))
    files = (C.c_char * 5)(*[b.file.encode()[:1] or b" " for b in c.B])
    names = (C.c_char * 100)(*b"".join(b.name.encode().ljust(20)[:20] for b in c.B))
    infos = (C.c_char * 10)(*b"".join(b.info.encode().ljust(2)[:2] for b in c.B))

    def arr(f):
        return (C.c_double * 5)(*[getattr(b, f) for b in c.B])
    thr = C.c_int32(0)
    lib.ref_build_bfield(C.c_int32(c.nB), (C.c_int32 * 3)(*c.n_B), files, names, infos,
                         arr("scale"), arr("B0"), arr("dL"), arr("x0"), arr("y0"), arr("z0"),
                         (C.c_double * 3)(*box), out.ctypes.data_as(C.c_void_p), C.byref(thr))
    return out, thr.value


BFIELD_CASES = {
    "gauss_y": [("n", "gauss", 1.0, 75e-4, "By", 0.5e-2, 1.1e-2)],
    "gauss_off": [("n", "gauss", 1.0, 75e-4, "By", -1.0, 1.1e-2)],
    "pg_z": [("n", "pg", 1.0, 3e-3, "Bz", 0.0, 0.0)],
    "cusp": [("n", "ext", 1.0, 875e-4, "By", 1.2e-2, 3.2e-2, 0.0, 0.6e-2)],
    "thruster": [("n", "thr", 1.0, 100e-4, "By", 0.53e-2, 0.7e-2)],
    "thruster_neg": [("n", "thr", 1.0, -100e-4, "By", 0.53e-2, 0.7e-2)],
    "mixed": [("s", "gauss", 1.0, 1.0, "By", 1.0, 0.0), ("n", "gauss", 1.0, 60e-4, "Bx", 2e-2, 1e-2),
              ("n", "ext", 2.0, 1e-2, "Bz", 1e-2, 1e-2, 0.0, 1e-2)],
}


@need_lib
@pytest.mark.parametrize("key", sorted(BFIELD_CASES))
@pytest.mark.parametrize("n_B", [(10, 6, 5), (1, 1, 1)])
def test_bfield_matches_reference(key, n_B):
    c = D.Conditions()
    recs = BFIELD_CASES[key]
    for b, r in zip(c.B, recs):
        b.file, b.name, b.scale, b.B0, b.info, b.dL, b.x0 = r[:7]
        if len(r) > 7:
            b.y0, b.z0 = r[7], r[8]
    c.nB = len(recs)
    c.flag_B = int(any(r[0] != "s" for r in recs))
    c.n_B = n_B
    box = (0.04, 0.03, 0.02)
    mine = D.build_bfield(c, *box)
    ref, thr = _ref_bfield(c, box)
    np.testing.assert_array_equal(mine.Bi, ref)
    assert mine.flag_thr == thr
    if key.startswith("thruster"):
        assert thr == 1


# ---------------------------------------------------------------------------
# step parameters, case assembly, writers
# ---------------------------------------------------------------------------

@need_ref
def test_reference_deck_params_and_case():
    d = D.load_deck(os.path.join(REF, "example_sans"), chemistry_dir=REF)
    c, p, b = d.conditions, d.params, d.boundary
    hmin = min(b.h)
    assert p.dt * p.vt0[0] / hmin == pytest.approx(c.kt, rel=1e-15)
    assert p.vt0[0] == pytest.approx(np.sqrt(2 * cfg.QE * c.Te / cfg.ME), rel=1e-15)
    for k, s in enumerate(d.chemistry.species):
        dens = c.n0 if s.charge != 0 else c.ngas
        assert p.Nm[k] * abs(c.np_cell) / (b.h[0] * b.h[1] * b.h[2]) == pytest.approx(dens, rel=1e-14)
    assert (p.nb_step_sort, p.nb_step_collisions) == (10, 10)
    assert p.nb_step_heating == (4 if c.kt >= 0.1 else 20 if c.kt >= 0.05 else 40)
    assert d.legacy_only() == []
    case = d.case(steps=10)
    assert case.Nm == p.Nm and case.dt == p.dt
    assert case.vt0[:2] == p.vt0[:2]
    np.testing.assert_array_equal(case.phi0(), b.u)
    assert case.flag_pbc == 1 and case.chemistry is not None


@need_ref
def test_legacy_only_features():
    """Injection is a legacy-driver feature: the modular step (and a case)
    runs the Thruster deck without it, with a warning, or refuses when strict."""
    d = D.load_deck(os.path.join(REF, "Thruster"))
    assert d.legacy_only() == ["particle injection"]
    with pytest.raises(D.DeckError, match="injection"):
        d.case(strict=True)
    with pytest.warns(UserWarning, match="injection"):
        case = d.case()
    assert case.flag_pbc == 1 and case.flag_die == 1 and list(case.dtype) == [0, 1, 1, 2, 3, 2]
    # the Hall-thruster field on its 128 x 256 x 96 map, the field grid being 256 x 512 x 192
    assert case.B_grid[0] == (128, 256, 96) and case.B_map.shape == (99, 259, 131, 4)


@need_ref
@pytest.mark.parametrize("name", DECKS_OK)
def test_conditions_writer_round_trip(name):
    c = D.read_conditions(os.path.join(REF, name, "conditions.inp"))
    back = D.read_conditions(D.write_conditions(c))
    for f in dataclasses.fields(c):
        a, b = getattr(c, f.name), getattr(back, f.name)
        if f.name == "B":
            for x, y in zip(a[:c.nB], b[:c.nB]):
                for g in dataclasses.fields(x):
                    u, v = getattr(x, g.name), getattr(y, g.name)
                    assert u == pytest.approx(v, rel=1e-15) if isinstance(u, float) else u == v
        elif isinstance(a, float):
            assert a == pytest.approx(b, rel=1e-15), f.name
        elif isinstance(a, list):
            assert a == pytest.approx(b, rel=1e-15), f.name
        else:
            assert a == b, f.name


def test_written_deck_loads(tmp_path):
    """A deck written by the writers loads back into a runnable case."""
    sp = cfg.argon_species()
    E = cfg.argon_energy_grid(40)
    reactions = [CH.Reaction("", (0, 2), (0, 2), 0.0, 0.0, 1, 1, E.copy(), np.full_like(E, 1e-19))]
    (tmp_path / "chem.dat").write_text(CH.write_reactions(sp, [1.0, 1.0, 1.0], reactions))
    c = D.Conditions(rname="chem.dat", Te=3.0, ng=3, eps=1e-8, omega=1.4, n=(16, 8, 8),
                     kt=0.3, n0=1e15, ngas=1e20, np_cell=4, ngrid=2, nsav=100, cnt_plt=1,
                     Pabs=-1.0, nu_h=-1.0, I_inj=-1.0)
    (tmp_path / "conditions.inp").write_text(D.write_conditions(c))
    (tmp_path / "geometry.inp").write_text(SLAB_GEO)
    (tmp_path / "boundary.inp").write_text(D.write_boundary_potentials([0.0, 4.0]))
    d = D.load_deck(str(tmp_path))
    assert d.legacy_only() == []
    case = d.case(steps=5)
    assert case.n == (16, 8, 8) and case.flag_pbc == 1
    assert [s.name for s in case.tracked] == ["e", "Ar+"]
    assert case.chemistry["ncol"] if "ncol" in case.chemistry else True
