"""The oracle against the reference's own Fortran routines.

oracle/ref.mk compiles the reference's MPI-free modules (mod_RNG,
mod_particles, mod_particle_sorting, mod_particleMover, mod_particleBC,
mod_chargeDeposition, mod_charge_weights, mod_electricField) with amdflang,
strict IEEE, single thread, plus a bind(C) bridge, into oracle/_ref.  Where
that library exists the restatement must agree with it bit for bit; the
routines that need MPI (density, Poisson, MCC) cannot be built here and are
pinned by tests/test_oracle.py instead."""
import ctypes as C
import os

import numpy as np
import pytest

from essaitemporaire_lepic_module_amd import config as cfg
from tests import helpers as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_ref", "libref_lepic.so")

pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="reference build (oracle/_ref) absent")


@pytest.fixture(scope="module")
def ref():
    L = C.CDLL(LIB)
    L.ref_ran2.restype = C.c_double
    return L


def _d(a):
    return a.ctypes.data_as(C.c_void_p)


def _n(n):
    return (C.c_int32 * 3)(*n), None


def _arrs(arr, count):
    out = [np.ascontiguousarray(np.array(a, np.float64)) for a in arr]
    return out


def test_ran2_matches_reference(ref, oracle):
    s = C.c_int32(123456)
    got = [ref.ref_ran2(C.byref(s)) for _ in range(1000)]
    want, seed = oracle.ran2(123456, 1000)
    np.testing.assert_array_equal(got, want)
    assert s.value == seed


def test_sort_matches_reference(ref, oracle, rng):
    n, h = H.box((7, 6, 5))
    N = 4000
    arr = _arrs(H.random_particles(rng, n, h, N, outside=0.03), N)
    dead = (rng.random(N) < 0.1).astype(np.int8)
    nn = (C.c_int32 * 3)(*n)
    hh = (C.c_double * 3)(*h)
    x, y, z, vx, vy, vz = (a.copy() for a in arr)
    dd = dead.copy()
    cell = np.zeros(N, np.int32)
    nc = n[0] * n[1] * n[2]
    cc, cs = np.zeros(nc, np.int32), np.zeros(nc, np.int32)
    ref.ref_sort(nn, hh, C.c_int64(N), _d(x), _d(y), _d(z), _d(vx), _d(vy), _d(vz), _d(dd), _d(cell),
                 _d(cc), _d(cs))
    p = oracle.Particles(N).set(*arr, dead=dead)
    occ, ocs = oracle.sort_particles(n, h, p)
    a = p.arrays()
    for k, v in zip(("x", "y", "z", "vx", "vy", "vz", "dead"), (x, y, z, vx, vy, vz, dd)):
        np.testing.assert_array_equal(a[k], v)
    np.testing.assert_array_equal(a["cell"] + 1, cell)          # reference ids are 1-based
    np.testing.assert_array_equal(occ, cc)
    np.testing.assert_array_equal(ocs + 1, cs)                  # 1-based starts


def test_move_electrostatic_matches_reference(ref, oracle, rng):
    n, h = H.box((6, 7, 5))
    N = 3000
    arr = _arrs(H.random_particles(rng, n, h, N, outside=0.03, vscale=4e5), N)
    E4 = H.random_E4(rng, n)
    E3 = np.ascontiguousarray(E4[..., :3])
    x, y, z, vx, vy, vz = (a.copy() for a in arr)
    ref.ref_move_es((C.c_int32 * 3)(*n), (C.c_double * 3)(*h), C.c_int64(N), _d(x), _d(y), _d(z),
                    _d(vx), _d(vy), _d(vz), _d(E3), C.c_double(-1.6e-19), C.c_double(9.1e-31),
                    C.c_double(5e-11))
    p = oracle.Particles(N).set(*arr)
    oracle.move_electrostatic(n, h, p, E4, -1.6e-19, 9.1e-31, 5e-11)
    a = p.arrays()
    for k, v in zip(("x", "y", "z", "vx", "vy", "vz"), (x, y, z, vx, vy, vz)):
        np.testing.assert_array_equal(a[k], v)


@pytest.mark.parametrize("bmaker,pbc", [(H.two_electrode_bcnd, 0), (H.periodic_yz_bcnd, 1)])
def test_apply_bc_matches_reference(ref, oracle, rng, bmaker, pbc):
    n, h = H.box((6, 6, 6))
    bc = np.ascontiguousarray(bmaker(n))
    ngrid = int(bc.max())
    N = 4000
    arr = _arrs(H.random_particles(rng, n, h, N, outside=0.08), N)
    dead = (rng.random(N) < 0.1).astype(np.int8)
    x, y, z, vx, vy, vz = (a.copy() for a in arr)
    dd = dead.copy()
    npart = C.c_int64(N)
    # the reference reads p_mac through a 1-based dummy: label L lands in slot L-1
    p_mac = np.zeros((ngrid + 1, 2))
    P_loss = C.c_double(0.0)
    ref.ref_apply_bc((C.c_int32 * 3)(*n), (C.c_double * 3)(*h), C.byref(npart), _d(x), _d(y), _d(z),
                     _d(vx), _d(vy), _d(vz), _d(dd), _d(bc), C.c_int32(pbc), C.c_int32(0),
                     C.c_int32(1), C.c_int32(-1), C.c_double(2.0), C.c_double(3.0), C.c_double(5.0),
                     C.c_int32(ngrid), _d(p_mac), C.byref(P_loss))
    p = oracle.Particles(N).set(*arr, dead=dead)
    lost, pmac_o, Ploss_o = oracle.apply_bc(n, h, p, bc, pbc, 0, 1, -1, 2.0, 3.0, 5.0, ngrid)
    m = npart.value
    assert p.n == m and lost == N - m
    a = p.arrays()
    for k, v in zip(("x", "y", "z", "vx", "vy", "vz", "dead"), (x, y, z, vx, vy, vz, dd)):
        np.testing.assert_array_equal(a[k], v[:m])
    np.testing.assert_array_equal(pmac_o.reshape(ngrid + 1, 2)[1:], p_mac[:ngrid])
    assert Ploss_o == P_loss.value


def test_deposit_and_kq_match_reference(ref, oracle, rng):
    n, h = H.box((6, 5, 7))
    bc = np.ascontiguousarray(H.two_electrode_bcnd(n))
    kq_r = np.zeros(bc.shape)
    ref.ref_build_kq((C.c_int32 * 3)(*n), _d(bc), _d(kq_r))
    kq_o = oracle.build_kq(n, bc)
    np.testing.assert_array_equal(kq_o, kq_r)
    N = 5000
    arr = _arrs(H.random_particles(rng, n, h, N, outside=0.01), N)
    dead = (rng.random(N) < 0.1).astype(np.int8)
    dep_r = np.zeros(bc.shape)
    ref.ref_deposit((C.c_int32 * 3)(*n), (C.c_double * 3)(*h), C.c_int64(N), _d(arr[0]), _d(arr[1]),
                    _d(arr[2]), _d(dead), _d(kq_r), C.c_double(1526.3), _d(dep_r))
    p = oracle.Particles(N).set(*arr, dead=dead)
    dep_o = oracle.deposit(n, h, p, kq_o, 1526.3)
    np.testing.assert_array_equal(dep_o, dep_r)


@pytest.mark.parametrize("bmaker", [cfg.grounded_box_bcnd, H.two_electrode_bcnd, H.periodic_yz_bcnd])
def test_efield_matches_reference(ref, oracle, rng, bmaker):
    n, h = H.box((7, 6, 8))
    bc = np.ascontiguousarray(bmaker(n))
    phi = rng.standard_normal(bc.shape)
    E3 = np.zeros(bc.shape + (3,))
    ref.ref_efield((C.c_int32 * 3)(*n), (C.c_double * 3)(*h), _d(phi), _d(E3), _d(bc))
    E4 = oracle.efield(n, h, phi, bc)
    np.testing.assert_array_equal(E4[..., :3], E3)


@pytest.mark.parametrize("pbc", [0, 1])
def test_apply_bc_dielectric_matches_reference(ref, oracle, rng, pbc):
    """Dielectric surface charging (flag_die = 1) on a Thruster-like
    geometry: x-z (type 2) and y-z (types 3, 4) dielectrics.  The reference
    reads the wall types through its 1-based dtype(:) dummy; the oracle must
    agree bit for bit on the surviving particles, the wall sums and both
    charge planes."""
    from essaitemporaire_lepic_module_amd import deck as D
    geo = D.read_geometry("2. 4. 1.5\n0. 1.25 0. 0.75 -2.75 -1.5 3\n0.74 1.25 0. -0.77 2.75 -1.5 4\n"
                          "0.77 0. 0. 2. -4. -1.5 5\n0. 1.25 0. 0. 2.75 1.5 -2\n2. 0. 0. 2. 4 1.5 -1\n")
    n = (16, 24, 10)
    b = D.generate_boundary(n, [0.0, 300.0, 0.0, 0.0, 0.0], 5, geo, 0.0015, 0.0075)
    h = b.h
    bc = np.ascontiguousarray(b.bcnd)
    # a table where every dielectric type appears under the shifted lookup
    dtype = np.array([0, 2, 3, 4, 2, 1], np.int32)
    ngrid = 5
    N = 6000
    arr = _arrs(H.random_particles(rng, n, h, N, outside=0.15), N)
    dead = (rng.random(N) < 0.05).astype(np.int8)
    x, y, z, vx, vy, vz = (a.copy() for a in arr)
    dd = dead.copy()
    npart = C.c_int64(N)
    p_mac = np.zeros((ngrid + 1, 2))
    P_loss = C.c_double(0.0)
    sxz = np.zeros((n[2] + 3, n[0] + 3))
    syz = np.zeros((n[2] + 3, n[1] + 3, 2))
    ref.ref_apply_bc_die((C.c_int32 * 3)(*n), (C.c_double * 3)(*h), C.byref(npart), _d(x), _d(y),
                         _d(z), _d(vx), _d(vy), _d(vz), _d(dd), _d(bc), C.c_int32(pbc), C.c_int32(0),
                         C.c_int32(1), C.c_int32(-1), C.c_double(-2.0), C.c_double(3.0),
                         C.c_double(5.0), C.c_int32(ngrid), _d(dtype), _d(sxz), _d(syz), _d(p_mac),
                         C.byref(P_loss))
    p = oracle.Particles(N).set(*arr, dead=dead)
    oxz, oyz = np.zeros_like(sxz), np.zeros_like(syz)
    lost, pmac_o, Ploss_o = oracle.apply_bc(n, h, p, bc, pbc, 0, 1, -1, -2.0, 3.0, 5.0, ngrid,
                                            dtype, oxz, oyz)
    m = npart.value
    assert p.n == m and lost == N - m
    a = p.arrays()
    for k, v in zip(("x", "y", "z", "vx", "vy", "vz"), (x, y, z, vx, vy, vz)):
        np.testing.assert_array_equal(a[k], v[:m])
    np.testing.assert_array_equal(pmac_o.reshape(ngrid + 1, 2)[1:], p_mac[:ngrid])
    assert Ploss_o == P_loss.value
    assert np.abs(sxz).sum() > 0 and np.abs(syz[..., 0]).sum() > 0 and np.abs(syz[..., 1]).sum() > 0
    np.testing.assert_array_equal(oxz, sxz)
    np.testing.assert_array_equal(oyz, syz)
