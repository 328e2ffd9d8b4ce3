#!/usr/bin/env python3
"""Writes tests/golden/reference_vectors.npz: inputs and outputs of the
reference's own Fortran routines, run from the library that oracle/ref.mk
builds into oracle/_ref (the reference's modules compiled in place with a
bind(C) bridge; strict IEEE, one thread).

Routines and the reference lines they run:
  ran2                          modules/core/mod_RNG.f90:28-45
  sort_particles_by_cell        modules/particles/mod_particle_sorting.f90:88-181
  move_particles_electrostatic  modules/particles/mod_particleMover.f90:72-156
  apply_particle_bc_legacy      modules/particles/mod_particleBC.f90:42-259
  build_kq                      modules/fields/mod_charge_weights.f90:9-33
  deposit_particle_set_to_np_thread  modules/fields/mod_chargeDeposition.f90:25-133
  calc_Efield_modular           modules/fields/mod_electricField.f90:9-164

Run in a container that holds the reference and has built oracle/_ref:
    make -C oracle -f ref.mk && python3 tests/golden/make_golden.py
The fixture is data only (arrays in, arrays out); it is what the tests read
where the reference is absent.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from essaitemporaire_lepic_module_amd import config as cfg  # noqa: E402
from tests import helpers as H  # noqa: E402

LIB = os.path.join(ROOT, "oracle", "_ref", "libref_lepic.so")
OUT = os.path.join(HERE, "reference_vectors.npz")


def _d(a):
    return a.ctypes.data_as(C.c_void_p)


def main():
    L = C.CDLL(LIB)
    L.ref_ran2.restype = C.c_double
    rng = np.random.default_rng(20261202)
    out = {}

    # ran2: 2000 draws from 123456
    s = C.c_int32(123456)
    out["ran2_draws"] = np.array([L.ref_ran2(C.byref(s)) for _ in range(2000)])
    out["ran2_seed_after"] = np.array([s.value], np.int32)

    # sort: 7x6x5 cells, 1500 particles (3 % outside the box), 10 % dead
    n, h = H.box((7, 6, 5))
    N = 1500
    arr = [np.ascontiguousarray(a) for a in H.random_particles(rng, n, h, N, outside=0.03)]
    dead = (rng.random(N) < 0.1).astype(np.int8)
    for k, a in zip(("x", "y", "z", "vx", "vy", "vz"), arr):
        out[f"sort_in_{k}"] = a.copy()
    out["sort_in_dead"] = dead.copy()
    x, y, z, vx, vy, vz = (a.copy() for a in arr)
    dd = dead.copy()
    cell = np.zeros(N, np.int32)
    nc = n[0] * n[1] * n[2]
    cc, cs = np.zeros(nc, np.int32), np.zeros(nc, np.int32)
    L.ref_sort((C.c_int32 * 3)(*n), (C.c_double * 3)(*h), C.c_int64(N), _d(x), _d(y), _d(z),
               _d(vx), _d(vy), _d(vz), _d(dd), _d(cell), _d(cc), _d(cs))
    for k, a in zip(("x", "y", "z", "vx", "vy", "vz", "dead", "cell1", "count", "start1"),
                    (x, y, z, vx, vy, vz, dd, cell, cc, cs)):
        out[f"sort_out_{k}"] = a
    out["sort_n"] = np.array(n, np.int32)

    # move_particles_electrostatic: 6x7x5 cells, 1500 particles, random E
    n, h = H.box((6, 7, 5))
    N = 1500
    arr = [np.ascontiguousarray(a) for a in H.random_particles(rng, n, h, N, outside=0.03, vscale=4e5)]
    E4 = H.random_E4(rng, n)
    E3 = np.ascontiguousarray(E4[..., :3])
    for k, a in zip(("x", "y", "z", "vx", "vy", "vz"), arr):
        out[f"move_in_{k}"] = a.copy()
    out["move_E4"] = E4
    x, y, z, vx, vy, vz = (a.copy() for a in arr)
    q, m, dt = -1.6e-19, 9.1e-31, 5e-11
    L.ref_move_es((C.c_int32 * 3)(*n), (C.c_double * 3)(*h), C.c_int64(N), _d(x), _d(y), _d(z),
                  _d(vx), _d(vy), _d(vz), _d(E3), C.c_double(q), C.c_double(m), C.c_double(dt))
    for k, a in zip(("x", "y", "z", "vx", "vy", "vz"), (x, y, z, vx, vy, vz)):
        out[f"move_out_{k}"] = a
    out["move_n"] = np.array(n, np.int32)
    out["move_qmdt"] = np.array([q, m, dt])

    # apply_particle_bc_legacy: two geometries on 6^3
    for tag, bmaker, pbc in (("bc0", H.two_electrode_bcnd, 0), ("bc1", H.periodic_yz_bcnd, 1)):
        n, h = H.box((6, 6, 6))
        bc = np.ascontiguousarray(bmaker(n))
        ngrid = int(bc.max())
        N = 1500
        arr = [np.ascontiguousarray(a) for a in H.random_particles(rng, n, h, N, outside=0.08)]
        dead = (rng.random(N) < 0.1).astype(np.int8)
        for k, a in zip(("x", "y", "z", "vx", "vy", "vz"), arr):
            out[f"{tag}_in_{k}"] = a.copy()
        out[f"{tag}_in_dead"] = dead.copy()
        out[f"{tag}_bcnd"] = bc
        x, y, z, vx, vy, vz = (a.copy() for a in arr)
        dd = dead.copy()
        npart = C.c_int64(N)
        p_mac = np.zeros((ngrid + 1, 2))
        P_loss = C.c_double(0.0)
        L.ref_apply_bc((C.c_int32 * 3)(*n), (C.c_double * 3)(*h), C.byref(npart), _d(x), _d(y),
                       _d(z), _d(vx), _d(vy), _d(vz), _d(dd), _d(bc), C.c_int32(pbc), C.c_int32(0),
                       C.c_int32(1), C.c_int32(-1), C.c_double(2.0), C.c_double(3.0),
                       C.c_double(5.0), C.c_int32(ngrid), _d(p_mac), C.byref(P_loss))
        mm = npart.value
        for k, a in zip(("x", "y", "z", "vx", "vy", "vz", "dead"), (x, y, z, vx, vy, vz, dd)):
            out[f"{tag}_out_{k}"] = a[:mm]
        # the reference reads p_mac through a 1-based dummy: label L lands in slot L-1
        out[f"{tag}_out_p_mac_slots"] = p_mac[:ngrid]
        out[f"{tag}_out_P_loss"] = np.array([P_loss.value])
        out[f"{tag}_params"] = np.array([pbc, 0, 1, -1, ngrid], np.int32)   # pbc nmn ptype tag ngrid
        out[f"{tag}_qmN"] = np.array([2.0, 3.0, 5.0])                     # qmacro mass Nm

    # build_kq + deposit: 6x5x7, 2000 particles, 10 % dead
    n, h = H.box((6, 5, 7))
    bc = np.ascontiguousarray(H.two_electrode_bcnd(n))
    kq = np.zeros(bc.shape)
    L.ref_build_kq((C.c_int32 * 3)(*n), _d(bc), _d(kq))
    N = 2000
    arr = [np.ascontiguousarray(a) for a in H.random_particles(rng, n, h, N, outside=0.01)]
    dead = (rng.random(N) < 0.1).astype(np.int8)
    dep = np.zeros(bc.shape)
    Nm = 1526.3
    L.ref_deposit((C.c_int32 * 3)(*n), (C.c_double * 3)(*h), C.c_int64(N), _d(arr[0]), _d(arr[1]),
                  _d(arr[2]), _d(dead), _d(kq), C.c_double(Nm), _d(dep))
    out["dep_bcnd"] = bc
    out["dep_kq"] = kq
    for k, a in zip(("x", "y", "z"), arr[:3]):
        out[f"dep_in_{k}"] = a
    out["dep_in_dead"] = dead
    out["dep_Nm"] = np.array([Nm])
    out["dep_out_np"] = dep

    # calc_Efield_modular: three geometries on 7x6x8
    for tag, bmaker in (("e0", cfg.grounded_box_bcnd), ("e1", H.two_electrode_bcnd),
                        ("e2", H.periodic_yz_bcnd)):
        n, h = H.box((7, 6, 8))
        bc = np.ascontiguousarray(bmaker(n))
        phi = rng.standard_normal(bc.shape)
        E3 = np.zeros(bc.shape + (3,))
        L.ref_efield((C.c_int32 * 3)(*n), (C.c_double * 3)(*h), _d(phi), _d(E3), _d(bc))
        out[f"{tag}_bcnd"] = bc
        out[f"{tag}_phi"] = phi
        out[f"{tag}_out_E"] = E3

    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}: {len(out)} arrays, {os.path.getsize(OUT)} bytes")


if __name__ == "__main__":
    main()
