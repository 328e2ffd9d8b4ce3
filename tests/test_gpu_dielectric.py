"""Dielectric surface charging on the device against the oracle: the wall BC
of the exact-order path (lepic_apply_bc), of the production in-place push
(k_push) and of the fused exact path, the potential update
(lepic_dielectric_phi), and whole steps of a run deck with dielectric walls.
Charge planes are summed by FP64 atomics (1e-12 relative); particles are
bit-identical."""
import dataclasses

import numpy as np
import pytest
import torch

from essaitemporaire_lepic_module_amd import config as cfg
from tests import helpers as H

pytestmark = pytest.mark.gpu

RTOL = 1e-12
THRUSTER_LIKE = ("2. 4. 1.5\n0. 1.25 0. 0.75 -2.75 -1.5 3\n0.74 1.25 0. -0.77 2.75 -1.5 4\n"
                 "0.77 0. 0. 2. -4. -1.5 5\n0. 1.25 0. 0. 2.75 1.5 -2\n2. 0. 0. 2. 4 1.5 -1\n")
DTYPE = np.array([0, 2, 3, 4, 2, 1], np.int32)     # every dielectric type under the shifted lookup


def _geometry(n):
    from essaitemporaire_lepic_module_amd import deck as D
    b = D.generate_boundary(n, [0.0, 300.0, 0.0, 0.0, 0.0], 5, D.read_geometry(THRUSTER_LIKE),
                            0.0015, 0.0075)
    return b.h, np.ascontiguousarray(b.bcnd)


def _planes(n, dev=None):
    sxz = np.zeros((n[2] + 3, n[0] + 3))
    syz = np.zeros((n[2] + 3, n[1] + 3, 2))
    if dev is None:
        return sxz, syz
    return (torch.zeros(sxz.shape, dtype=torch.float64, device=dev),
            torch.zeros(syz.shape, dtype=torch.float64, device=dev))


def _close(g, o):
    np.testing.assert_allclose(g.cpu().numpy(), o, rtol=0, atol=RTOL * max(np.abs(o).max(), 1e-300))


@pytest.mark.parametrize("pbc", [0, 1])
def test_apply_bc_dielectric(gpu, oracle, rng, pbc):
    from essaitemporaire_lepic_module_amd import particles as P
    n = (16, 24, 10)
    h, bc = _geometry(n)
    count = 40000
    arr = H.random_particles(rng, n, h, count, outside=0.12)
    dead = (rng.random(count) < 0.03).astype(np.int8)
    ps = P.ParticleSet(count + 16, gpu)
    ps.from_host(*arr, dead=dead)
    p_mac = torch.zeros((6, 2), dtype=torch.float64, device=gpu)
    P_loss = torch.zeros(1, dtype=torch.float64, device=gpu)
    gxz, gyz = _planes(n, gpu)
    dtg = torch.as_tensor(DTYPE, device=gpu)
    lost = P.apply_particle_bc_legacy(ps, n, h, torch.as_tensor(bc, device=gpu), pbc, 0, 1, -1, 1,
                                      -2.4e-16, p_mac, 9.1e-31, P_loss, 1500.0, dtg, gxz, gyz)
    po = oracle.Particles(count + 16)
    po.set(*arr, dead=dead)
    oxz, oyz = _planes(n)
    lost_o, pmac_o, Ploss_o = oracle.apply_bc(n, h, po, bc, pbc, 0, 1, -1, -2.4e-16, 9.1e-31,
                                              1500.0, 5, DTYPE, oxz, oyz)
    assert lost == lost_o and ps.n == po.n
    g, o = ps.to_host(), po.arrays()
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        np.testing.assert_array_equal(g[k], o[k])
    assert np.abs(oxz).sum() > 0 and np.abs(oyz).sum() > 0
    _close(gxz, oxz)
    _close(gyz, oyz)
    np.testing.assert_allclose(p_mac.cpu().numpy().reshape(-1), pmac_o, rtol=RTOL, atol=0)


@pytest.mark.parametrize("inplace", [True, False])
def test_push_dielectric(gpu, oracle, rng, inplace):
    """The fused particle phases charge the dielectrics as mover + BC do."""
    from essaitemporaire_lepic_module_amd import particles as P
    n = (16, 24, 10)
    h, bc = _geometry(n)
    kq = cfg.build_kq(bc)
    count = 30000
    arr = H.random_particles(rng, n, h, count, outside=0.02, vscale=3e6)
    E4 = H.random_E4(rng, n)
    q, m, dt, Nm = -1.60217646e-19, 9.10938188e-31, 4e-11, 1526.0
    po = oracle.Particles(count + 16)
    po.set(*arr)
    oracle.sort_particles(n, h, po)
    start = po.arrays()
    ps = P.ParticleSet(count + 16, gpu)
    ps.from_host(*(start[k] for k in ("x", "y", "z", "vx", "vy", "vz")))
    gxz, gyz = _planes(n, gpu)
    dtg = torch.as_tensor(DTYPE, device=gpu)
    bp = P.bc_params(n, h, 1, 0, 1, -1, Nm * q, m, Nm, 5, 1, dtg, gxz, gyz)
    bf = P.bfield(0)
    p_mac = torch.zeros((6, 2), dtype=torch.float64, device=gpu)
    P_loss = torch.zeros(1, dtype=torch.float64, device=gpu)
    npg = torch.zeros(cfg.node_shape(n), dtype=torch.float64, device=gpu)
    bcg = torch.as_tensor(bc, device=gpu)
    kqg = torch.as_tensor(kq, device=gpu)
    E4g = torch.as_tensor(E4, device=gpu)
    if inplace:
        P.push_bc_deposit_inplace(ps, n, h, E4g, bf, q, m, dt, bcg, bp, kqg, npg, p_mac, P_loss)
    else:
        P.push_bc_deposit(ps, n, h, E4g, bf, q, m, dt, bcg, bp, kqg, npg, p_mac, P_loss)
    oracle.move_electrostatic(n, h, po, E4, q, m, dt)
    oxz, oyz = _planes(n)
    oracle.apply_bc(n, h, po, bc, 1, 0, 1, -1, Nm * q, m, Nm, 5, DTYPE, oxz, oyz)
    g = ps.to_host()
    keep = g["dead"] == 0
    o = po.arrays()
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        np.testing.assert_array_equal(g[k][keep], o[k])
    assert np.abs(oxz).sum() + np.abs(oyz).sum() > 0
    _close(gxz, oxz)
    _close(gyz, oyz)


def test_dielectric_phi(gpu, oracle, rng):
    from essaitemporaire_lepic_module_amd import fields as F
    n = (16, 24, 10)
    h, bc = _geometry(n)
    dtype = np.array([0, 1, 1, 2, 3, 4], np.int32)
    sxz = rng.normal(0, 1e-13, (2, n[2] + 3, n[0] + 3))
    syz = rng.normal(0, 1e-13, (2, n[2] + 3, n[1] + 3, 2))
    phi = rng.normal(0, 1, cfg.node_shape(n))
    Ca = 1e-6 * h[0] * h[2]
    pg = torch.as_tensor(phi.copy(), device=gpu)
    F.apply_dielectric_bc_to_phi(n, h, torch.as_tensor(bc, device=gpu),
                                 torch.as_tensor(dtype, device=gpu), 5,
                                 torch.as_tensor(sxz, device=gpu), torch.as_tensor(syz, device=gpu),
                                 Ca, pg)
    oracle.dielectric_phi(n, h, bc, dtype, 5, sxz, syz, Ca, phi)
    np.testing.assert_array_equal(pg.cpu().numpy(), phi)


def test_dielectric_deck_steps_match_oracle(gpu, oracle, tmp_path):
    """A written deck with x-z and y-z dielectrics, periodic along z, runs the
    same steps on the GPU as in the oracle (exact-order path, ran2 MCC)."""
    from essaitemporaire_lepic_module_amd import chemistry as CH
    from essaitemporaire_lepic_module_amd import deck as D
    from essaitemporaire_lepic_module_amd import particles as P
    from tests.test_gpu_step import _compare, _start
    sp = cfg.argon_species()
    E = cfg.argon_energy_grid(60)
    reactions = [CH.Reaction("", (0, 2), (0, 2), 0.0, 0.0, 1, 1, E.copy(), np.full_like(E, 2e-19))]
    (tmp_path / "chem.dat").write_text(CH.write_reactions(sp, [1.0, 1.0, 1.0], reactions))
    c = D.Conditions(rname="chem.dat", Te=3.0, ng=3, eps=1e-8, omega=1.4, n=(8, 16, 8), kt=0.3,
                     n0=1e15, ngas=1e20, np_cell=6, ngrid=5, nsav=100, cnt_plt=1, Pabs=-1.0,
                     nu_h=-1.0, I_inj=-1.0, Ca=1e-6)
    (tmp_path / "conditions.inp").write_text(D.write_conditions(c))
    (tmp_path / "geometry.inp").write_text(THRUSTER_LIKE)
    (tmp_path / "boundary.inp").write_text(D.write_boundary_potentials([0.0, 30.0, 0.0, 0.0, 0.0]))
    d = D.load_deck(str(tmp_path))
    case = dataclasses.replace(d.case(steps=6), nb_step_sort=4, nb_step_collisions=2)
    assert case.flag_die == 1 and case.flag_pbc == 1
    ref, sim = _start(oracle, case, P, lazy=False)
    for istep in range(1, 7):
        sim.advance_one_step(istep)
        ref.advance_one_step(istep)
        torch.cuda.synchronize()
        _compare(sim, ref, exact=istep == 1, step=istep)
        _close(sim.sum_q_xz, ref.sum_q_xz)
        _close(sim.sum_q_yz, ref.sum_q_yz)
    assert float(sim.sum_q_xz.abs().sum() + sim.sum_q_yz.abs().sum()) > 0.0
