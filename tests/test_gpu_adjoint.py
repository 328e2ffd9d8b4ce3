"""The adjoint kernels (lepic_adj_*) against torch autograd of the same
operators written independently in tests/torch_pic.py (FP64, CPU), and the
whole backward (adjoint.DiffPIC) against the pure-torch 16^3 run and against
central finite differences of the GPU forward.

Tolerances: single operators 1e-10 relative to the largest entry (FP64
rounding of differently ordered sums); the Poisson adjoint and whole-run
gradients 1e-6 (multigrid solves to a residual tolerance, dense solve in
torch); finite differences as stated in each test (the run is piecewise
smooth: the cloud-in-cell weights have kinks at cell faces, so a central
difference carries an error that shrinks with the step)."""
import ctypes as C
import math

import numpy as np
import pytest
import torch

from essaitemporaire_lepic_module_amd import _native
from essaitemporaire_lepic_module_amd import config as cfg
from tests.torch_pic import TorchPIC

pytestmark = pytest.mark.gpu

QE = cfg.QE
ME = cfg.ME
MI = 39.95 * cfg.AMU
DT = 5.605424054411371e-11


def _ctx():
    from essaitemporaire_lepic_module_amd.device import context
    return context("cuda:0")


def _box(n):
    h = (3.125e-4,) * 3
    return tuple(n), h, cfg.biased_wall_bcnd(tuple(n), 2)


def _particles(rng, n, h, N, vscale):
    L = np.array([n[d] * h[d] for d in range(3)])
    pos = (0.05 + 0.9 * rng.random((N, 3))) * L
    vel = vscale * rng.standard_normal((N, 3))
    return pos, vel


def _dev_state(pos, vel, dead=None):
    N = len(pos)
    d = {k: torch.as_tensor(np.ascontiguousarray(a), device="cuda:0")
         for k, a in zip(("x", "y", "z", "vx", "vy", "vz"), (*pos.T, *vel.T))}
    d["dead"] = torch.zeros(N + 4, dtype=torch.int8, device="cuda:0")
    if dead is not None:
        d["dead"][:N] = torch.as_tensor(dead)
    p = _native.PSet()
    for k in ("x", "y", "z", "vx", "vy", "vz", "dead"):
        setattr(p, k, d[k].data_ptr())
    p.cell = None
    p.n = N
    p.cap = N
    return p, d


def _adj_set(arrs):
    t = {k: torch.as_tensor(np.ascontiguousarray(a), device="cuda:0").clone()
         for k, a in zip(("lx", "ly", "lz", "lvx", "lvy", "lvz"), arrs)}
    s = _native.AdjSet()
    for k, v in t.items():
        setattr(s, k, v.data_ptr())
    s.n = len(arrs[0])
    return s, t


def _close(a, b, rtol):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(np.abs(b).max(), 1e-300)
    assert np.abs(a - b).max() <= rtol * scale, (np.abs(a - b).max(), scale)


@pytest.mark.parametrize("B0", [None, 0.005, 0.05])
def test_adj_push_matches_autograd(gpu, rng, B0):
    """Boris + gather + position update of one step: dL/dx, dL/dv, dL/dE and
    dL/dB0 of L = <lx', x'> + <lv', v'>."""
    n, h, bc = _box((8, 8, 8))
    N = 3000
    q, m = -QE, ME
    pos, vel = _particles(rng, n, h, N, 1e6)
    shape = cfg.node_shape(n)
    E = 1e4 * rng.standard_normal(shape + (3,))
    E4 = np.zeros(shape + (4,))
    E4[..., :3] = E
    lx, lv = rng.standard_normal((N, 3)), rng.standard_normal((N, 3)) * 1e-7
    ctx = _ctx()
    g = _native.grid(n, h)
    p, keep = _dev_state(pos, vel)
    a, lam = _adj_set((*lx.T, *lv.T))
    lamE = torch.zeros(shape + (4,), dtype=torch.float64, device=gpu)
    lamB = torch.zeros(3, dtype=torch.float64, device=gpu)
    E4d = torch.as_tensor(E4, device=gpu)
    Bp = None if B0 is None else (C.c_double * 3)(0.0, 0.0, B0)
    ctx.check(ctx.lib.lepic_adj_push(ctx.h, C.byref(g), C.byref(p), C.byref(a), E4d.data_ptr(), Bp,
                                     q, m, DT, lamE.data_ptr(), lamB.data_ptr()))
    torch.cuda.synchronize()
    model = TorchPIC(n, h, bc, 2, [(q, m, 1.0)], DT)
    x = torch.tensor(pos, requires_grad=True)
    v = torch.tensor(vel, requires_grad=True)
    Et = torch.tensor(E, requires_grad=True)
    Bt = torch.tensor(0.0 if B0 is None else B0, dtype=torch.float64, requires_grad=True)
    Ep = model.gather(Et, x)
    if B0 is None:
        vn = v + 2.0 * (DT * q / (2.0 * m)) * Ep
    else:
        vn = model.boris(v, Ep, Bt, q, m)
    xn = x + DT * vn
    L = (torch.tensor(lx) * xn).sum() + (torch.tensor(lv) * vn).sum()
    gx, gv, gE, gB = torch.autograd.grad(L, [x, v, Et, Bt], allow_unused=True)
    got_x = np.stack([lam[k].cpu().numpy() for k in ("lx", "ly", "lz")], 1)
    got_v = np.stack([lam[k].cpu().numpy() for k in ("lvx", "lvy", "lvz")], 1)
    _close(got_x, gx.numpy(), 1e-10)
    _close(got_v, gv.numpy(), 1e-10)
    _close(lamE.cpu().numpy()[..., :3], gE.numpy(), 1e-10)
    if B0 is not None:
        assert abs(lamB[2].item() - gB.item()) <= 1e-10 * abs(gB.item())


@pytest.mark.parametrize("magnetized", [True, False])
def test_adj_push_map_matches_autograd(gpu, rng, magnetized):
    """The push with a B map (gathered with the particle's weights, gated at
    its lower node): dL/dx (with the B gather's own position term), dL/dv and
    dL/dE of L = <lx', x'> + <lv', v'>, against autograd."""
    n, h, bc = _box((8, 8, 8))
    N = 3000
    q, m = -QE, ME
    pos, vel = _particles(rng, n, h, N, 1e6)
    shape = cfg.node_shape(n)
    E = 1e4 * rng.standard_normal(shape + (3,))
    E4 = np.zeros(shape + (4,))
    E4[..., :3] = E
    B4 = cfg.radial_bfield_map(n, h, B0=0.05, z_peak=0.5, width=0.3)
    B4[:, :, 2:4] = 0.0                       # i = 2, 3: below the gate
    lx, lv = rng.standard_normal((N, 3)), rng.standard_normal((N, 3)) * 1e-7
    ctx = _ctx()
    g = _native.grid(n, h)
    p, keep = _dev_state(pos, vel)
    a, lam = _adj_set((*lx.T, *lv.T))
    lamE = torch.zeros(shape + (4,), dtype=torch.float64, device=gpu)
    E4d = torch.as_tensor(E4, device=gpu)
    B4d = torch.as_tensor(B4, device=gpu)
    ctx.check(ctx.lib.lepic_adj_push_map(ctx.h, C.byref(g), C.byref(p), C.byref(a),
                                         E4d.data_ptr(), B4d.data_ptr(), int(magnetized), q, m,
                                         DT, lamE.data_ptr()))
    torch.cuda.synchronize()
    model = TorchPIC(n, h, bc, 2, [(q, m, 1.0)], DT)
    x = torch.tensor(pos, requires_grad=True)
    v = torch.tensor(vel, requires_grad=True)
    Et = torch.tensor(E, requires_grad=True)
    Ep = model.gather(Et, x)
    Bp, use = model.gather_b(torch.tensor(B4), x)
    assert 0 < int(use.sum()) < N
    if not magnetized:
        use = torch.zeros_like(use)
    vn = model.boris_map(v, Ep, Bp, use, q, m)
    xn = x + DT * vn
    L = (torch.tensor(lx) * xn).sum() + (torch.tensor(lv) * vn).sum()
    gx, gv, gE = torch.autograd.grad(L, [x, v, Et])
    got_x = np.stack([lam[k].cpu().numpy() for k in ("lx", "ly", "lz")], 1)
    got_v = np.stack([lam[k].cpu().numpy() for k in ("lvx", "lvy", "lvz")], 1)
    _close(got_x, gx.numpy(), 1e-10)
    _close(got_v, gv.numpy(), 1e-10)
    _close(lamE.cpu().numpy()[..., :3], gE.numpy(), 1e-10)


def test_adj_push_skips_holes(gpu, rng):
    n, h, bc = _box((8, 8, 8))
    N = 500
    pos, vel = _particles(rng, n, h, N, 1e5)
    dead = (np.arange(N) % 3 == 0).astype(np.int8) * 2
    E4 = torch.as_tensor(np.random.default_rng(1).standard_normal(cfg.node_shape(n) + (4,)),
                         device=gpu)
    ctx = _ctx()
    g = _native.grid(n, h)
    p, keep = _dev_state(pos, vel, dead)
    ones = np.ones(N)
    a, lam = _adj_set((ones, ones, ones, ones, ones, ones))
    lamE = torch.zeros(cfg.node_shape(n) + (4,), dtype=torch.float64, device=gpu)
    ctx.check(ctx.lib.lepic_adj_push(ctx.h, C.byref(g), C.byref(p), C.byref(a), E4.data_ptr(), None,
                                     -QE, ME, DT, lamE.data_ptr(), None))
    for k in lam:
        np.testing.assert_array_equal(lam[k].cpu().numpy()[dead != 0], 1.0)


@pytest.mark.parametrize("geometry", ["block", "thin_gaps"])
def test_adj_efield_is_the_transpose(gpu, rng, geometry):
    """<lamE, D phi> = <D^T lamE, phi> with D = lepic_efield, and D^T lamE
    equal to torch autograd of the independent E restatement; with an
    electrode block inside, and with plates one plasma node apart (chains of
    wall-pass reads)."""
    from essaitemporaire_lepic_module_amd import fields as F
    from tests import helpers as H
    n, h, bc = _box((12, 10, 10))
    if geometry == "block":
        bc[4:6, 4:7, 5:8] = 3
    else:
        bc = H.thin_gap_bcnd(n)
    shape = cfg.node_shape(n)
    phi = rng.standard_normal(shape)
    lamE = rng.standard_normal(shape + (4,))
    ctx = _ctx()
    g = _native.grid(n, h)
    bcd = torch.as_tensor(bc, device=gpu)
    phid = torch.as_tensor(phi, device=gpu)
    E4 = torch.zeros(shape + (4,), dtype=torch.float64, device=gpu)
    F.calc_Efield_modular(n, h, phid, E4, bcd)
    lamEd = torch.as_tensor(lamE, device=gpu)
    lamphi = torch.zeros(shape, dtype=torch.float64, device=gpu)
    ctx.check(ctx.lib.lepic_adj_efield(ctx.h, C.byref(g), bcd.data_ptr(), lamEd.data_ptr(),
                                       lamphi.data_ptr()))
    lhs = float((lamEd[..., :3] * E4[..., :3]).sum())
    rhs = float((lamphi * phid).sum())
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + float((lamEd.abs() * E4.abs()).sum()))
    model = TorchPIC(n, h, bc, 2, [(-QE, ME, 1.0)], DT)
    pt = torch.tensor(phi.reshape(-1), requires_grad=True)
    Et = model.efield(pt)
    _close(E4.cpu().numpy()[..., :3], Et.detach().numpy(), 1e-13)
    (gphi,) = torch.autograd.grad((torch.tensor(lamE[..., :3]) * Et).sum(), [pt])
    _close(lamphi.cpu().numpy().reshape(-1), gphi.numpy(), 1e-12)


def test_adj_deposit_matches_autograd(gpu, rng):
    n, h, bc = _box((8, 8, 8))
    N = 4000
    q, Nm = QE, 3.8e3
    pos, vel = _particles(rng, n, h, N, 1e3)
    shape = cfg.node_shape(n)
    lamb = rng.standard_normal(shape)
    kq = cfg.build_kq(bc)
    ctx = _ctx()
    g = _native.grid(n, h)
    p, keep = _dev_state(pos, vel)
    z = np.zeros(N)
    a, lam = _adj_set((z, z, z, z, z, z))
    kqd = torch.as_tensor(kq, device=gpu)
    lambd = torch.as_tensor(lamb, device=gpu)   # both held while the kernel runs
    ctx.check(ctx.lib.lepic_adj_deposit(ctx.h, C.byref(g), C.byref(p), C.byref(a),
                                        kqd.data_ptr(), lambd.data_ptr(), q, Nm))
    torch.cuda.synchronize()
    model = TorchPIC(n, h, bc, 2, [(q, MI, Nm)], DT)
    x = torch.tensor(pos, requires_grad=True)
    rho = -q * model.deposit(x, torch.ones(N, dtype=torch.bool), Nm)
    (gx,) = torch.autograd.grad((torch.tensor(lamb.reshape(-1)) * rho).sum(), [x])
    got = np.stack([lam[k].cpu().numpy() for k in ("lx", "ly", "lz")], 1)
    _close(got, gx.numpy(), 1e-11)


def test_plane_flux_value_and_gradient(gpu, rng):
    n, h, bc = _box((8, 8, 8))
    N = 20000
    q, Nm = QE, 3.8e3
    pos, vel = _particles(rng, n, h, N, 1e3)
    ctx = _ctx()
    g = _native.grid(n, h)
    p, keep = _dev_state(pos, vel)
    z = np.zeros(N)
    a, lam = _adj_set((z, z, z, z, z, z))
    area = n[1] * h[1] * n[2] * h[2]
    coef = q * Nm / (area * h[0])
    acc = torch.zeros(1, dtype=torch.float64, device=gpu)
    ctx.check(ctx.lib.lepic_plane_flux(ctx.h, C.byref(g), C.byref(p), n[0], coef, acc.data_ptr(),
                                       C.byref(a), 1.0))
    model = TorchPIC(n, h, bc, 2, [(q, MI, Nm)], DT)
    x = torch.tensor(pos, requires_grad=True)
    v = torch.tensor(vel, requires_grad=True)
    f = model.plane_flux(x, v, torch.ones(N, dtype=torch.bool), q, Nm)
    gx, gv = torch.autograd.grad(f, [x, v])
    assert abs(acc.item() - f.item()) <= 1e-12 * abs(f.item())
    _close(lam["lx"].cpu().numpy(), gx[:, 0].numpy(), 1e-12)
    _close(lam["lvx"].cpu().numpy(), gv[:, 0].numpy(), 1e-12)


@pytest.mark.parametrize("axis,plane,sign", [(1, 3, 1.0), (2, 2, -1.0)])
def test_plane_flux_axis_value_and_gradient(gpu, rng, axis, plane, sign):
    """The loss plane normal to y or z (the annular channel's anode current)."""
    n, h, bc = _box((8, 7, 6))
    N = 20000
    q, Nm = QE, 3.8e3
    pos, vel = _particles(rng, n, h, N, 1e3)
    ctx = _ctx()
    g = _native.grid(n, h)
    p, keep = _dev_state(pos, vel)
    z = np.zeros(N)
    a, lam = _adj_set((z, z, z, z, z, z))
    coef = sign * q * Nm / h[axis]
    acc = torch.zeros(1, dtype=torch.float64, device=gpu)
    ctx.check(ctx.lib.lepic_plane_flux_axis(ctx.h, C.byref(g), C.byref(p), axis, plane, coef,
                                            acc.data_ptr(), C.byref(a), 1.0))
    model = TorchPIC(n, h, bc, 2, [(q, MI, Nm)], DT)
    x = torch.tensor(pos, requires_grad=True)
    v = torch.tensor(vel, requires_grad=True)
    f = model.plane_flux(x, v, torch.ones(N, dtype=torch.bool), q, Nm, (axis, plane, sign, 1.0))
    gx, gv = torch.autograd.grad(f, [x, v])
    assert f.item() != 0.0
    assert abs(acc.item() - f.item()) <= 1e-12 * abs(f.item())
    kx, kv = ("lx", "ly", "lz")[axis], ("lvx", "lvy", "lvz")[axis]
    _close(lam[kx].cpu().numpy(), gx[:, axis].numpy(), 1e-12)
    _close(lam[kv].cpu().numpy(), gv[:, axis].numpy(), 1e-12)
    for k in ("lx", "ly", "lz", "lvx", "lvy", "lvz"):
        if k not in (kx, kv):
            assert not lam[k].any()
    with pytest.raises(_native.LepicError):
        ctx.check(ctx.lib.lepic_plane_flux_axis(ctx.h, C.byref(g), C.byref(p), axis,
                                                n[axis] + 2, coef, acc.data_ptr(), None, 1.0))


def test_poisson_adjoint_matches_dense(gpu, rng):
    """mu = A^-1 dL/dphi by the forward's multigrid and dL/dV of the biased
    electrode, against autograd through a dense solve of the same operator."""
    from essaitemporaire_lepic_module_amd import fields as F
    n, h, bc = _box((16, 16, 16))
    shape = cfg.node_shape(n)
    lamphi = rng.standard_normal(shape)
    ctx = _ctx()
    g = _native.grid(n, h)
    bcd = torch.as_tensor(bc, device=gpu)
    lp = torch.as_tensor(lamphi, device=gpu)
    rhs = torch.zeros(shape, dtype=torch.float64, device=gpu)
    mu = torch.zeros(shape, dtype=torch.float64, device=gpu)
    sc = 1.0 / float(lp.abs().max())
    ctx.check(ctx.lib.lepic_adj_poisson_rhs(ctx.h, C.byref(g), bcd.data_ptr(), lp.data_ptr(), sc,
                                            rhs.data_ptr()))
    F.solve_poisson_legacy(mu, bcd, rhs, h, n, 20000, 1e-13, 1.4, 4, 0, 0)
    mu.mul_(1.0 / sc)
    lamV = torch.zeros(1, dtype=torch.float64, device=gpu)
    ctx.check(ctx.lib.lepic_adj_bias(ctx.h, C.byref(g), bcd.data_ptr(), 2, lp.data_ptr(),
                                     mu.data_ptr(), cfg.EPS0, lamV.data_ptr()))
    model = TorchPIC(n, h, bc, 2, [(QE, MI, 1.0)], DT)
    b = torch.tensor(rng.standard_normal(int(np.prod(shape))) * 1e-6, requires_grad=True)
    V = torch.tensor(20.0, dtype=torch.float64, requires_grad=True)
    phi = model.solve(b, V)
    gb, gV = torch.autograd.grad((torch.tensor(lamphi.reshape(-1)) * phi).sum(), [b, V])
    plasma = model.plasma.reshape(-1).numpy()
    _close(mu.cpu().numpy().reshape(-1)[plasma], gb.numpy()[plasma], 1e-8)
    assert abs(lamV.item() - gV.item()) <= 1e-8 * abs(gV.item())


def _initial(rng, case):
    n, h = case.n, case.h
    N = case.np_cell * n[0] * n[1] * n[2]
    L = np.array([n[d] * h[d] for d in range(3)])
    pos = rng.random((N, 3)) * L
    parts = []
    for s in case.tracked:
        vt = cfg.thermal_speed(s.T_eV, s.mass)
        vel = rng.standard_normal((N, 3)) * vt / math.sqrt(2.0)
        parts.append({"x": pos[:, 0], "y": pos[:, 1], "z": pos[:, 2],
                      "vx": vel[:, 0], "vy": vel[:, 1], "vz": vel[:, 2]})
    return parts


@pytest.mark.parametrize("steps,checkpoint,resort", [(12, 5, 0), (12, 5, 5), (12, 4, 5)])
def test_gradient_matches_pure_torch_16cube(gpu, steps, checkpoint, resort):
    """The whole backward at 16^3 x 4 ppc against autograd of the pure-torch
    run from the same particles, with the wall events the GPU forward saw;
    with the particles re-sorted by cell at steps 6 and 11 (at checkpoints,
    and inside the segments of 4 steps)."""
    from essaitemporaire_lepic_module_amd.adjoint import DiffPIC
    rng = np.random.default_rng(4)
    case = cfg.cfg4_differentiable(n=16, np_cell=4, steps=steps, mcc_every=0)
    parts = _initial(rng, case)
    sim = DiffPIC(case, parts, "cuda:0", checkpoint=checkpoint, resort=resort)
    L = sim.forward(20.0, 0.005, steps, keep=True, record_deaths=True)
    dV, dB = sim.backward()
    sched = [torch.as_tensor(sim.by_input(s, d)) for s, d in enumerate(sim.deaths)]
    sp = [(s.charge, s.mass, nm) for s, nm in zip(case.tracked, case.Nm)]
    model = TorchPIC(case.n, case.h, case.bcnd, 2, sp, case.dt)
    xs = [torch.tensor(np.stack([p["x"], p["y"], p["z"]], 1)) for p in parts]
    vs = [torch.tensor(np.stack([p["vx"], p["vy"], p["vz"]], 1)) for p in parts]
    V = torch.tensor(20.0, dtype=torch.float64, requires_grad=True)
    B0 = torch.tensor(0.005, dtype=torch.float64, requires_grad=True)
    Lt, _ = model.run(xs, vs, V, B0, steps, schedule=sched)
    gV, gB = torch.autograd.grad(Lt, [V, B0])
    assert abs(L - Lt.item()) <= 1e-8 * abs(Lt.item())
    assert abs(dV - gV.item()) <= 1e-6 * abs(gV.item()), (dV, gV.item())
    assert abs(dB - gB.item()) <= 1e-6 * abs(gB.item()), (dB, gB.item())


@pytest.mark.parametrize("resort", [0, 5])
def test_checkpoint_interval_leaves_gradient_unchanged(gpu, resort):
    """Checkpoints every 1, 4 and 12 steps give one gradient; MCC at step 10
    (recomputed segments replay the recorded collisions); with resort 5 the
    particles are re-sorted at steps 6 and 11, at checkpoints or inside the
    segments (the backward replays the recorded slot orders)."""
    from essaitemporaire_lepic_module_amd.adjoint import DiffPIC
    rng = np.random.default_rng(5)
    case = cfg.cfg4_differentiable(n=16, np_cell=4, steps=12)
    parts = _initial(rng, case)
    out = []
    for K in (1, 4, 12):
        sim = DiffPIC(case, parts, "cuda:0", checkpoint=K, resort=resort)
        sim.forward(20.0, 0.005, 12)
        assert sum(int(ev["idx"].numel()) for ev in sim.mcc_events[10]) > 0
        out.append(sim.backward())
    for dV, dB in out[1:]:
        assert abs(dV - out[0][0]) <= 1e-9 * abs(out[0][0])
        assert abs(dB - out[0][1]) <= 1e-9 * abs(out[0][1])


def test_gradient_matches_central_differences(gpu):
    """32^3 x 8 ppc, 20 steps, MCC at steps 10 and 20, particles re-sorted by
    cell at steps 9 and 17: dL/dV and dL/dB0 against central differences (steps 1e-3 and 1e-4 relative) of the run with
    the nominal wall events and collisions replayed (the map the adjoint
    differentiates).  Central differences of the free
    run, whose wall events move with the parameters, are printed beside
    them: their difference from the adjoint is the events' own share of the
    derivative, which the frozen-event gradient leaves out."""
    from essaitemporaire_lepic_module_amd.adjoint import DiffPIC, PicLoss
    rng = np.random.default_rng(6)
    steps = 20
    case = cfg.cfg4_differentiable(n=32, np_cell=8, steps=steps)
    parts = _initial(rng, case)
    sim = DiffPIC(case, parts, "cuda:0", checkpoint=10, resort=8)   # re-sorts at steps 9, 17
    V = torch.tensor(20.0, dtype=torch.float64, requires_grad=True)
    B0 = torch.tensor(0.005, dtype=torch.float64, requires_grad=True)
    L = PicLoss.apply(V, B0, sim, steps)
    L.backward()
    dV, dB = V.grad.item(), B0.grad.item()
    sim.forward(20.0, 0.005, steps, keep=False, record_deaths=True)
    deaths = [d.clone() for d in sim.deaths]
    events, sorts = sim.mcc_events, sim.sorts
    ncoll = sum(int(ev["idx"].numel()) for evs in events.values() for ev in evs)
    nprod = sum(ev["m1"] - ev["m0"] for evs in events.values() for ev in evs)
    print(f"collided {ncoll} products {nprod}")
    assert ncoll > 0
    # The frozen-event map is piecewise smooth: a particle crossing a cell
    # plane between the two runs puts a kink of its deposit weights inside
    # the difference, an error that shrinks with the step (for dL/dB0 it was
    # 1e-3 relative at a step of 1e-3).
    hV, hB = 20.0 * 1e-3, 0.005 * 1e-4

    def fd(frozen):
        f = lambda a, b: sim.forward(a, b, steps, keep=False, frozen=frozen,
                                     frozen_mcc=events if frozen is not None else None,
                                     frozen_sorts=sorts if frozen is not None else None)
        return ((f(20.0 + hV, 0.005) - f(20.0 - hV, 0.005)) / (2 * hV),
                (f(20.0, 0.005 + hB) - f(20.0, 0.005 - hB)) / (2 * hB))

    fV, fB = fd(deaths)
    gV, gB = fd(None)
    print(f"dL/dV adjoint {dV:.10e} frozen fd {fV:.10e} free fd {gV:.10e}")
    print(f"dL/dB0 adjoint {dB:.10e} frozen fd {fB:.10e} free fd {gB:.10e}")
    assert abs(dV - fV) <= 1e-4 * abs(fV)
    assert abs(dB - fB) <= 3e-4 * abs(fB)


def _initial_channel(rng, case):
    """np_cell particles per box cell drawn uniformly, kept where the cell
    has a plasma node among its corners (the annular channel's cells)."""
    n, h = case.n, case.h
    pl = np.asarray(case.bcnd) == -1
    load = np.zeros((n[2], n[1], n[0]), bool)
    for dk in (1, 2):
        for dj in (1, 2):
            for di in (1, 2):
                load |= pl[dk:dk + n[2], dj:dj + n[1], di:di + n[0]]
    N = case.np_cell * n[0] * n[1] * n[2]
    L = np.array([n[d] * h[d] for d in range(3)])
    pos = rng.random((N, 3)) * L
    c = np.minimum((pos / np.array(h)).astype(np.int64), np.array(n) - 1)
    pos = pos[load[c[:, 2], c[:, 1], c[:, 0]]]
    parts = []
    for s in case.tracked:
        vt = cfg.thermal_speed(s.T_eV, s.mass)
        vel = rng.standard_normal((len(pos), 3)) * vt / math.sqrt(2.0)
        parts.append({"x": pos[:, 0], "y": pos[:, 1], "z": pos[:, 2],
                      "vx": vel[:, 0], "vy": vel[:, 1], "vz": vel[:, 2]})
    return parts


def test_annular_map_gradient_matches_pure_torch(gpu):
    """The combined run's geometry at 16^3 x 4 ppc, 10 steps, no MCC: the
    annular channel, the anode (label 2) at 300 V, the radial Br(z) map on
    both species, the loss the ion current into the anode through z = h3.
    dL/dV against autograd of the pure-torch run from the same particles
    with the wall events the GPU forward saw."""
    from essaitemporaire_lepic_module_amd.adjoint import DiffPIC
    rng = np.random.default_rng(8)
    steps = 10
    case = cfg.cfg5_combined(n=16, np_cell=4, steps=steps)
    case.chemistry = None
    parts = _initial_channel(rng, case)
    sim = DiffPIC(case, parts, "cuda:0", checkpoint=4)
    L = sim.forward(300.0, None, steps, keep=True, record_deaths=True)
    dV, dB = sim.backward()
    assert dB == 0.0
    sched = [torch.as_tensor(sim.by_input(s, d)) for s, d in enumerate(sim.deaths)]
    assert sum(int((d > 0).sum()) for d in sched) > 0
    sp = [(s.charge, s.mass, nm) for s, nm in zip(case.tracked, case.Nm)]
    model = TorchPIC(case.n, case.h, case.bcnd, 2, sp, case.dt)
    xs = [torch.tensor(np.stack([p["x"], p["y"], p["z"]], 1)) for p in parts]
    vs = [torch.tensor(np.stack([p["vx"], p["vy"], p["vz"]], 1)) for p in parts]
    V = torch.tensor(300.0, dtype=torch.float64, requires_grad=True)
    Lt, _ = model.run(xs, vs, V, None, steps, schedule=sched, B4=torch.tensor(case.B_map),
                      flux=case.flux_plane)
    gV, = torch.autograd.grad(Lt, [V])
    print(f"L {L:.10e} torch {Lt.item():.10e} dL/dV {dV:.10e} torch {gV.item():.10e}")
    assert gV.item() != 0.0
    assert abs(L - Lt.item()) <= 1e-8 * abs(Lt.item())
    assert abs(dV - gV.item()) <= 1e-6 * abs(gV.item()), (dV, gV.item())


def test_annular_map_gradient_matches_central_differences(gpu):
    """The combined run's geometry at 32^3 x 8 ppc, 20 steps, MCC at steps 10
    and 20: dL/dV of the anode current against central differences (step 1e-4
    relative) of the run with the nominal wall events and collisions
    replayed."""
    from essaitemporaire_lepic_module_amd.adjoint import DiffPIC, PicLoss
    rng = np.random.default_rng(9)
    steps = 20
    case = cfg.cfg5_combined(n=32, np_cell=8, steps=steps)
    parts = _initial_channel(rng, case)
    sim = DiffPIC(case, parts, "cuda:0", checkpoint=10, resort=8)   # re-sorts at steps 9, 17
    V = torch.tensor(300.0, dtype=torch.float64, requires_grad=True)
    L = PicLoss.apply(V, None, sim, steps)
    L.backward()
    dV = V.grad.item()
    sim.forward(300.0, None, steps, keep=False, record_deaths=True)
    deaths = [d.clone() for d in sim.deaths]
    events, sorts = sim.mcc_events, sim.sorts
    ncoll = sum(int(ev["idx"].numel()) for evs in events.values() for ev in evs)
    assert ncoll > 0
    hV = 300.0 * 1e-4
    f = lambda a, fr: sim.forward(a, None, steps, keep=False, frozen=fr,
                                  frozen_mcc=events if fr is not None else None,
                                  frozen_sorts=sorts if fr is not None else None)
    fV = (f(300.0 + hV, deaths) - f(300.0 - hV, deaths)) / (2 * hV)
    gV = (f(300.0 + hV, None) - f(300.0 - hV, None)) / (2 * hV)
    print(f"collided {ncoll}; dL/dV adjoint {dV:.10e} frozen fd {fV:.10e} free fd {gV:.10e}")
    assert fV != 0.0
    assert abs(dV - fV) <= 1e-4 * abs(fV)


def test_pcg_solves_leave_gradient_unchanged(gpu):
    """The differentiable run with its forward and adjoint Poisson solves by
    the preconditioned conjugate gradients (to the same 1e-10): the loss and
    dL/dV of the annular channel agree with the multigrid run's to the
    solves' tolerance."""
    from essaitemporaire_lepic_module_amd.adjoint import DiffPIC
    rng = np.random.default_rng(8)
    case = cfg.cfg5_combined(n=16, np_cell=4, steps=10)
    parts = _initial_channel(rng, case)
    out = []
    for poisson in ("fp64", "pcg"):
        sim = DiffPIC(case, parts, "cuda:0", checkpoint=4, poisson=poisson)
        L = sim.forward(300.0, None, 10, keep=True)
        dV, _ = sim.backward()
        out.append((L, dV))
    print(out)
    assert abs(out[1][0] - out[0][0]) <= 1e-7 * abs(out[0][0])
    assert abs(out[1][1] - out[0][1]) <= 1e-6 * abs(out[0][1])
