"""The pure-torch reference of the differentiable step (tests/torch_pic.py)
held to the oracle operator by operator (so the gradient it gives is the
gradient of the reference's step), and its autograd gradient to central
differences of its own forward with the wall events frozen."""
import math

import numpy as np
import pytest
import torch

from essaitemporaire_lepic_module_amd import config as cfg
from tests.torch_pic import TorchPIC

QE = cfg.QE


def _setup(n=(12, 10, 10)):
    """The biased box with an electrode block inside, two plasma nodes or
    more from every other electrode."""
    h = (3.125e-4,) * 3
    bc = cfg.biased_wall_bcnd(n, 2)
    bc[4:6, 4:7, 5:8] = 3
    return n, h, bc


def _thin_gaps(n=(12, 10, 10)):
    """Electrode plates one plasma node apart along z (k = 1 wall, plates at
    k = 3 and 5): the wall pass's chains of second-neighbour reads."""
    h = (3.125e-4,) * 3
    bc = cfg.biased_wall_bcnd(n, 2)
    bc[3, 3:7, 3:8] = 3
    bc[5, 4:6, 4:7] = 4
    return n, h, bc


@pytest.mark.parametrize("geometry", ["block", "thin_gaps"])
def test_efield_equals_oracle(oracle, rng, geometry):
    n, h, bc = _setup() if geometry == "block" else _thin_gaps()
    phi = rng.standard_normal(bc.shape)
    model = TorchPIC(n, h, bc, 2, [(-QE, cfg.ME, 1.0)], 1e-11)
    E = model.efield(torch.tensor(phi.reshape(-1))).numpy()
    Eo = oracle.efield(n, h, phi, bc)
    np.testing.assert_allclose(E, Eo[..., :3], rtol=0, atol=1e-13 * np.abs(Eo).max())


def test_deposit_and_push_equal_oracle(oracle, rng):
    n, h, bc = _setup()
    L = np.array([n[d] * h[d] for d in range(3)])
    N = 3000
    pos = (0.05 + 0.9 * rng.random((N, 3))) * L
    vel = 1e6 * rng.standard_normal((N, 3))
    q, m, Nm, dt, Bz = -QE, cfg.ME, 3.8e3, 2e-11, 0.01
    model = TorchPIC(n, h, bc, 2, [(q, m, Nm)], dt)
    kq = oracle.build_kq(n, bc)
    p = oracle.Particles(N).set(*pos.T, *vel.T)
    np_o = oracle.deposit(n, h, p, kq, Nm)
    np_t = model.deposit(torch.tensor(pos), torch.ones(N, dtype=torch.bool), Nm).numpy()
    np.testing.assert_allclose(np_t, np_o.reshape(-1), rtol=1e-13, atol=1e-13 * np_o.max())
    E4 = np.zeros(bc.shape + (4,))
    E4[..., :3] = 1e4 * rng.standard_normal(bc.shape + (3,))
    oracle.move_boris(n, h, p, E4, 1, (0.0, 0.0, Bz), None, q, m, dt)
    a = p.arrays()
    Ep = model.gather(torch.tensor(E4[..., :3]), torch.tensor(pos))
    v = model.boris(torch.tensor(vel), Ep, torch.tensor(Bz, dtype=torch.float64), q, m)
    x = torch.tensor(pos) + dt * v
    for d, k in enumerate(("vx", "vy", "vz")):
        np.testing.assert_allclose(v[:, d].numpy(), a[k], rtol=1e-13, atol=1e-6)
    for d, k in enumerate(("x", "y", "z")):
        np.testing.assert_allclose(x[:, d].numpy(), a[k], rtol=1e-14, atol=0)


@pytest.mark.parametrize("magnetized", [True, False])
def test_map_push_equals_oracle(oracle, rng, magnetized):
    """Gather of a B map, gated at the lower node, and Boris with the
    per-particle field: the radial Br(z) map with a slab of nodes below the
    1e-5 T gate, against the oracle's b_mode 2 push."""
    n, h, bc = _setup()
    L = np.array([n[d] * h[d] for d in range(3)])
    N = 3000
    pos = (0.05 + 0.9 * rng.random((N, 3))) * L
    vel = 1e6 * rng.standard_normal((N, 3))
    q, m, dt = -QE, cfg.ME, 2e-11
    B4 = cfg.radial_bfield_map(n, h, B0=0.05, z_peak=0.5, width=0.3)
    B4[:, :, 3:6] = 0.0                       # i = 3..5: no field
    model = TorchPIC(n, h, bc, 2, [(q, m, 1.0)], dt)
    p = oracle.Particles(N).set(*pos.T, *vel.T)
    E4 = np.zeros(bc.shape + (4,))
    E4[..., :3] = 1e4 * rng.standard_normal(bc.shape + (3,))
    oracle.move_boris(n, h, p, E4, 2, None, B4, q, m, dt, magnetized=magnetized)
    a = p.arrays()
    xt = torch.tensor(pos)
    Ep = model.gather(torch.tensor(E4[..., :3]), xt)
    Bp, use = model.gather_b(torch.tensor(B4), xt)
    assert 0 < int(use.sum()) < N
    if not magnetized:
        use = torch.zeros_like(use)
    v = model.boris_map(torch.tensor(vel), Ep, Bp, use, q, m)
    for d, k in enumerate(("vx", "vy", "vz")):
        np.testing.assert_allclose(v[:, d].numpy(), a[k], rtol=1e-13, atol=1e-6)


def test_solve_equals_oracle(oracle, rng):
    n, h, bc = _setup((12, 10, 10))
    model = TorchPIC(n, h, bc, 2, [(-QE, cfg.ME, 1.0)], 1e-11)
    rho = np.where(bc == -1, 1e-6 * rng.standard_normal(bc.shape), 0.0)
    phi_t = model.solve(torch.tensor(rho.reshape(-1)), torch.tensor(15.0)).numpy()
    phi_o = np.where(bc == 2, 15.0, 0.0)
    st, _ = oracle.pdesolver(n, h, phi_o, rho.copy(), bc, eps=1e-13, ng=2)
    assert st == 0
    np.testing.assert_allclose(phi_t, phi_o.reshape(-1), rtol=0, atol=1e-9 * 15.0)


def test_autograd_gradient_matches_frozen_differences():
    """16^3 x 2 ppc, 8 steps: dL/dV and dL/dB0 against central differences
    with the wall events of the nominal run (h = 1e-4 relative)."""
    n = (16, 16, 16)
    h = (3.125e-4,) * 3
    bc = cfg.biased_wall_bcnd(n, 2)
    n0, ppc = 1e15, 2
    Nm = n0 * h[0] * h[1] * h[2] / ppc
    sp = [(-QE, cfg.ME, Nm), (QE, 39.95 * cfg.AMU, Nm)]
    dt = 0.1 / cfg.omega_pe(n0)
    model = TorchPIC(n, h, bc, 2, sp, dt)
    rng = np.random.default_rng(0)
    N = ppc * 16 ** 3
    L = 16 * h[0]
    pos = rng.random((N, 3)) * L
    vt = [cfg.thermal_speed(3.0, cfg.ME), cfg.thermal_speed(0.03, 39.95 * cfg.AMU)]
    xs = [torch.tensor(pos), torch.tensor(pos)]
    vs = [torch.tensor(rng.standard_normal((N, 3)) * v / math.sqrt(2)) for v in vt]
    V = torch.tensor(20.0, dtype=torch.float64, requires_grad=True)
    B0 = torch.tensor(0.005, dtype=torch.float64, requires_grad=True)
    L0, sched = model.run(xs, vs, V, B0, 8)
    gV, gB = torch.autograd.grad(L0, [V, B0])
    assert sum(int((s > 0).sum()) for s in sched) > 0     # walls absorbed some
    with torch.no_grad():
        def f(a, b):
            return model.run(xs, vs, torch.tensor(a, dtype=torch.float64),
                             torch.tensor(b, dtype=torch.float64), 8, schedule=sched)[0].item()
        r = 1e-4
        fV = (f(20.0 * (1 + r), 0.005) - f(20.0 * (1 - r), 0.005)) / (2 * 20.0 * r)
        fB = (f(20.0, 0.005 * (1 + r)) - f(20.0, 0.005 * (1 - r))) / (2 * 0.005 * r)
    # The frozen-event map is piecewise smooth (a particle crossing a cell
    # plane puts a kink in its deposit weights), so the differences' error
    # shrinks with the step: 2e-4 of dL/dB0 at r = 1e-3, 2e-7 at r = 1e-4.
    assert abs(gV.item() - fV) <= 1e-5 * abs(fV)
    assert abs(gB.item() - fB) <= 1e-5 * abs(fB)


def test_annular_map_gradient_matches_frozen_differences():
    """The combined run's geometry at 16^3 x 4 ppc, 10 steps: the annular
    channel, the anode at 300 V, the radial Br(z) map, the loss the ion
    current into the anode through z = h3; dL/dV against central differences
    with the nominal wall events (h = 1e-4 relative)."""
    case = cfg.cfg5_combined(n=16, np_cell=4, steps=10)
    n, h = case.n, case.h
    rng = np.random.default_rng(8)
    N = 4 * 16 ** 3
    L = np.array([n[d] * h[d] for d in range(3)])
    pos = rng.random((N, 3)) * L
    c = np.minimum((pos / np.array(h)).astype(np.int64), 15)
    near = (case.bcnd[1:17, 1:17, 1:17] == -1) | (case.bcnd[2:18, 2:18, 2:18] == -1)
    pos = pos[near[c[:, 2], c[:, 1], c[:, 0]]]
    sp = [(s.charge, s.mass, nm) for s, nm in zip(case.tracked, case.Nm)]
    model = TorchPIC(n, h, case.bcnd, 2, sp, case.dt)
    vt = [cfg.thermal_speed(s.T_eV, s.mass) for s in case.tracked]
    xs = [torch.tensor(pos), torch.tensor(pos)]
    vs = [torch.tensor(rng.standard_normal(pos.shape) * v / math.sqrt(2)) for v in vt]
    B4 = torch.tensor(case.B_map)
    V = torch.tensor(300.0, dtype=torch.float64, requires_grad=True)
    L0, sched = model.run(xs, vs, V, None, 10, B4=B4, flux=case.flux_plane)
    gV, = torch.autograd.grad(L0, [V])
    assert sum(int((s > 0).sum()) for s in sched) > 0
    with torch.no_grad():
        def f(a):
            return model.run(xs, vs, torch.tensor(a, dtype=torch.float64), None, 10,
                             schedule=sched, B4=B4, flux=case.flux_plane)[0].item()
        r = 1e-4
        fV = (f(300.0 * (1 + r)) - f(300.0 * (1 - r))) / (2 * 300.0 * r)
    assert gV.item() != 0.0
    assert abs(gV.item() - fV) <= 1e-5 * abs(fV)
