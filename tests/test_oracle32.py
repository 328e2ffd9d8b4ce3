"""The oracle's FP32 particle phase pinned against its FP64 phase (itself
pinned against the reference's modules): the same particles, fields and
walls through one step in each precision, each result within a tolerance
that FP32 rounding fixes, and exact where the precision form promises
exactness (conversion round trips, charge conservation, stable sort)."""
import math

import numpy as np
import pytest

from essaitemporaire_lepic_module_amd import config as cfg
from tests import helpers as H

QE = 1.60217646e-19
ME = 9.10938188e-31


def _state(oracle, rng, n, h, count, vscale=3e5, margin=0.0):
    arr = H.random_particles(rng, n, h, count, vscale=vscale, margin=margin)
    p64 = oracle.Particles(count + 16).set(*arr)
    p32, out = oracle.from64(n, h, p64)
    assert out == 0
    # the FP64 state the FP32 one represents exactly
    return p32, oracle.to64(n, h, p32)


def test_conversion_exact(oracle, rng):
    """to64(from64(p)) moves no particle by more than half an FP32 ulp of its
    offset, keeps cells, and from64(to64(q)) gives q back bit for bit;
    particles outside [0, L) and dead ones become holes."""
    n, h = H.box((9, 7, 6))
    count = 5000
    arr = H.random_particles(rng, n, h, count, outside=0.05)
    dead = (rng.random(count) < 0.05).astype(np.int8)
    p64 = oracle.Particles(count).set(*arr, dead=dead)
    q, out = oracle.from64(n, h, p64)
    L = np.array(n) * np.array(h)
    inside = np.all([(arr[d] >= 0) & (arr[d] < L[d]) for d in range(3)], 0)
    assert out == int((inside == 0)[dead == 0].sum())
    a = q.arrays()
    live = a["cell"] >= 0
    np.testing.assert_array_equal(live, inside & (dead == 0))
    back = oracle.to64(n, h, q).arrays()
    for d, k in enumerate("xyz"):
        err = np.abs(back[k][live] - arr[d][live]) / h[d]
        assert err.max() <= 2.0 ** -24 + 1e-15
    for k in ("vx", "vy", "vz"):
        np.testing.assert_array_equal(a[k], arr[3 + "xyz".index(k[1])].astype(np.float32))
    q2, out2 = oracle.from64(n, h, oracle.to64(n, h, q))
    assert out2 == 0
    for k in q.arrays():
        np.testing.assert_array_equal(q2.arrays()[k], a[k])
    # offsets in [0, 1)
    for k in ("fx", "fy", "fz"):
        assert (a[k][live] >= 0).all() and (a[k][live] < 1).all()


@pytest.mark.parametrize("bmode", [0, 1])
def test_push32_against_fp64(oracle, rng, bmode):
    """One push from the same state: FP32 velocities within 2e-6 relative of
    the FP64 step's, positions within 1e-6 of a cell plus 2e-6 of the
    distance moved, the same particles lost with the same wall diagnostics
    (1e-6), and survivors in the FP64 step's cell unless they sit within
    1e-5 cell of a face."""
    n, h = H.box((10, 8, 7))
    bc = cfg.grounded_box_bcnd(n)
    count = 20000
    p32, p64 = _state(oracle, rng, n, h, count, vscale=8e5)
    E4 = H.random_E4(rng, n, 3e3)
    E4f = oracle.round4(E4)
    E4r = E4f.astype(np.float64)   # both precisions see the same (FP32-representable) field
    B = (0.0, 0.0, 0.01)
    q, m, dt = -QE, ME, 5.6e-11
    lost32, esc, pm32, Pl32 = oracle.push32(n, h, p32, E4f, bmode, B, None, q, m, dt, bc, 0, 0, 1,
                                            -1, -QE * 10, ME, 10.0, 1)
    assert esc == 0
    oracle.move_boris(n, h, p64, E4r, bmode, B, None, q, m, dt)
    lost64, pm64, Pl64 = oracle.apply_bc(n, h, p64, bc, 0, 0, 1, -1, -QE * 10, ME, 10.0, 1)
    assert lost32 == lost64
    a = p32.arrays()
    live = a["cell"] >= 0
    assert live.sum() == count - lost64
    back = oracle.to64(n, h, p32).arrays()
    # survivors in order: the FP64 BC compacts them stably
    s64 = p64.arrays()
    for k in ("vx", "vy", "vz"):
        np.testing.assert_allclose(back[k][live], s64[k], rtol=2e-6, atol=2e-6 * 8e5)
    for d, k in enumerate("xyz"):
        err = np.abs(back[k][live] - s64[k]) / h[d]
        step = np.abs(s64["v" + k]) * dt / h[d]
        assert np.all(err <= 1e-6 + 2e-6 * step)
    c = [np.floor(s64[k] / h[d]).astype(np.int64) for d, k in enumerate("xyz")]
    cell64 = c[0] + n[0] * (c[1] + n[1] * c[2])
    near = np.zeros(live.sum(), bool)
    for d, k in enumerate("xyz"):
        f = s64[k] / h[d]
        near |= np.abs(f - np.rint(f)) < 1e-5
    same = a["cell"][live] == cell64
    assert np.all(same | near) and same.mean() > 0.999
    np.testing.assert_allclose(pm32, pm64, rtol=1e-6, atol=1e-30)
    np.testing.assert_allclose(Pl32, Pl64, rtol=1e-6)


def test_deposit32_exact_and_close(oracle, rng):
    """The FP32 set's fixed-point deposit: the same for any particle order,
    exactly N full weights on the grid, and within 1e-6 relative (plus the
    fixed-point rounding) of the FP64 deposit of the same positions."""
    n, h = H.box((8, 8, 8))
    bc = cfg.grounded_box_bcnd(n)
    kq = oracle.build_kq(n, bc)
    count = 30000
    p32, p64 = _state(oracle, rng, n, h, count)
    Nm = 1526.0
    k1 = Nm / (h[0] * h[1] * h[2])
    S = 22 - math.ceil(math.log2(k1))
    dep = oracle.deposit32(n, h, p32, Nm, S)
    perm = rng.permutation(count)
    a = p32.arrays()
    q = oracle.Particles32(count).set(*(a[k][perm] for k in oracle.FIELDS32))
    np.testing.assert_array_equal(oracle.deposit32(n, h, q, Nm, S), dep)
    U = round(k1 * 2.0 ** S)
    assert int(dep.sum()) == count * U
    rho32 = oracle.fixed_density(n, dep, kq, S)
    rho64 = oracle.deposit(n, h, p64, kq, Nm)
    scale = np.abs(rho64).max()
    np.testing.assert_allclose(rho32, rho64, rtol=0, atol=1e-6 * scale)


def test_sort32_stable(oracle, rng):
    """Stable by cell, holes dropped: the same as a stable numpy argsort."""
    n, h = H.box((6, 5, 4))
    count = 8000
    p32, _ = _state(oracle, rng, n, h, count)
    a = p32.arrays()
    a["cell"][rng.random(count) < 0.1] = -1
    p32.set(*(a[k] for k in oracle.FIELDS32))
    s = oracle.sort32(n, h, p32).arrays()
    keep = a["cell"] >= 0
    order = np.argsort(np.where(keep, a["cell"], 1 << 30), kind="stable")[: keep.sum()]
    for k in oracle.FIELDS32:
        np.testing.assert_array_equal(s[k], a[k][order])


def test_fixed_split32_exact(oracle, rng):
    """The FP32 split of a full weight of at most 2^22 units: 8 integers that
    sum to U exactly, each within 1.5 units of U times its trilinear weight;
    checked against exact rational arithmetic of the same rounding steps."""
    import ctypes as C
    lib = oracle.lib32()
    lib.lo32_fixed_split.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float,
                                     C.POINTER(C.c_int32)]
    a = (C.c_int32 * 8)()
    U = np.concatenate([[1.0, 2.0 ** 22, 2.0 ** 22 - 1], rng.integers(1, 2 ** 22, 2000)])
    u = rng.random((U.size, 3)).astype(np.float32)
    u[:5] = [[0, 0, 0], [0.5, 0.5, 0.5], [np.float32(1) - np.float32(2 ** -24)] * 3,
             [0.25, 0.75, 0.125], [0.999, 0.001, 0.5]]
    for Ui, (ux, uy, uz) in zip(U, u):
        lib.lo32_fixed_split(float(Ui), float(ux), float(uy), float(uz), a)
        parts = np.array(a[:], np.int64)
        assert parts.sum() == int(Ui)
        wx, wy, wz = 1 - float(ux), 1 - float(uy), 1 - float(uz)
        w = np.array([wx * wy * wz, (1 - wx) * wy * wz, (1 - wx) * (1 - wy) * wz,
                      wx * (1 - wy) * wz, wx * wy * (1 - wz), (1 - wx) * wy * (1 - wz),
                      (1 - wx) * (1 - wy) * (1 - wz), wx * (1 - wy) * (1 - wz)])
        assert np.all(np.abs(parts - Ui * w) <= 1.5 + 1e-9 * Ui)
