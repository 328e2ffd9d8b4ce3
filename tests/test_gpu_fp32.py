"""GPU parity of the FP32 particle phase against the oracle's FP32
restatement (itself pinned against the FP64 phase in test_oracle32.py):
cells, offsets, velocities, loss counts and the fixed-point density are
bit-identical; wall diagnostics (FP64 atomics) agree to 1e-12."""
import math

import numpy as np
import pytest
import torch

from essaitemporaire_lepic_module_amd import config as cfg
from tests import helpers as H

pytestmark = pytest.mark.gpu

QE = 1.60217646e-19
ME = 9.10938188e-31
AMU = 1.66053886e-27


def _bits(n, h, Nm):
    return 21 - math.ceil(math.log2(abs(Nm / (h[0] * h[1] * h[2]))))


def _pair(oracle, P32, rng, n, h, count, vscale=5e5, outside=0.0, dead=None):
    """The same FP32 set on the host (oracle) and on the device."""
    arr = H.random_particles(rng, n, h, count, vscale=vscale, outside=outside)
    p64 = oracle.Particles(count + 16).set(*arr, dead=dead)
    q, _ = oracle.from64(n, h, p64)
    a = q.arrays()
    g = P32.ParticleSet32(count + 16, "cuda:0").from_host(*(a[k] for k in oracle.FIELDS32))
    return q, g


def _same32(g, o):
    a, b = g.to_host(), o.arrays()
    for k in b:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_conversions_bitexact(gpu, oracle, rng):
    """FP64 -> FP32 on the device equals the oracle's (cells, offsets,
    velocities, outside count), and FP32 -> FP64 back as well."""
    from essaitemporaire_lepic_module_amd import particles as P
    from essaitemporaire_lepic_module_amd import particles32 as P32
    n, h = H.box((11, 9, 7))
    count = 20000
    arr = H.random_particles(rng, n, h, count, outside=0.03)
    dead = (rng.random(count) < 0.04).astype(np.int8)
    ps = P.ParticleSet(count + 16, gpu).from_host(*arr, dead=dead)
    g, out = P32.ParticleSet32.from_pset64(ps, n, h)
    q, oout = oracle.from64(n, h, oracle.Particles(count + 16).set(*arr, dead=dead))
    assert out == oout > 0
    _same32(g, q)
    back = g.to_pset64(n, h).to_host()
    ob = oracle.to64(n, h, q).arrays()
    for k in ("x", "y", "z", "vx", "vy", "vz", "dead"):
        np.testing.assert_array_equal(back[k], ob[k], err_msg=k)


@pytest.mark.parametrize("case", ["box_b0", "box_b1", "periodic_b1", "neumann_map"])
def test_push32_bitexact(gpu, oracle, rng, case):
    """The FP32 push + wall BC + fixed-point deposit against the oracle: every
    particle's cell, offsets and velocity bit for bit, the same losses, the
    same int64 density, wall diagnostics to 1e-12."""
    from essaitemporaire_lepic_module_amd import particles as P
    from essaitemporaire_lepic_module_amd import particles32 as P32
    n, h = H.box((12, 10, 8))
    flag_pbc = flag_nmn = 0
    bmode, B, B4 = {"box_b0": 0, "box_b1": 1, "periodic_b1": 1, "neumann_map": 2}[case], \
        (0.002, -0.001, 0.01), None
    if case == "periodic_b1":
        bc = H.periodic_yz_bcnd(n)
        flag_pbc = 1
    elif case == "neumann_map":
        bc = H.two_electrode_bcnd(n)
        bc[2:n[2] + 1, 2:n[1] + 1, 0:2] = -2    # a Neumann x = 0 face: reflection
        flag_nmn = 1
        B4 = np.zeros(cfg.node_shape(n) + (4,))
        B4[..., :3] = 0.01 * rng.standard_normal(cfg.node_shape(n) + (3,))
        B4[..., 3] = np.sqrt((B4[..., :3] ** 2).sum(-1))
        B4[rng.random(cfg.node_shape(n)) < 0.2, 3] = 0.0    # gated nodes
    else:
        bc = cfg.grounded_box_bcnd(n)
    ngrid = int(bc.max())
    count = 40000
    q, g = _pair(oracle, P32, rng, n, h, count, vscale=1.2e6)
    # sorted input, as in production, plus some holes
    q = oracle.sort32(n, h, q)
    a = q.arrays()
    a["cell"][rng.random(q.n) < 0.02] = -1
    q.set(*(a[k] for k in oracle.FIELDS32))
    g = P32.ParticleSet32(q.n + 16, gpu).from_host(*(a[k] for k in oracle.FIELDS32))
    E4 = H.random_E4(rng, n, 5e3)
    E4f = oracle.round4(E4)
    B4f = oracle.round4(B4) if B4 is not None else None
    qe, m, dt, Nm = -QE, ME, 5.6e-11, 1526.0
    S = _bits(n, h, Nm)
    dep_o = np.zeros(cfg.node_shape(n), np.int64)
    lost, esc, pm, Pl = oracle.push32(n, h, q, E4f, bmode, B, B4f, qe, m, dt, bc, flag_pbc,
                                      flag_nmn, 1, -1, Nm * qe, m, Nm, ngrid, S, dep_o)
    assert lost > 0
    Eg = torch.as_tensor(E4f, device=gpu)
    B4g = torch.as_tensor(B4f, device=gpu) if B4f is not None else None
    bf = P32.bfield32(bmode, B, B4g)
    bp = P.bc_params(n, h, flag_pbc, flag_nmn, 1, -1, Nm * qe, m, Nm, ngrid)
    dep = torch.zeros(cfg.node_shape(n), dtype=torch.int64, device=gpu)
    p_mac = torch.zeros(2 * (ngrid + 1), dtype=torch.float64, device=gpu)
    P_loss = torch.zeros(1, dtype=torch.float64, device=gpu)
    ctr = torch.zeros(2, dtype=torch.int64, device=gpu)
    P32.push32(g, n, h, Eg, bf, qe, m, dt, torch.as_tensor(bc, device=gpu), bp, S, dep, p_mac,
               P_loss, ctr[0:1], ctr[1:2])
    torch.cuda.synchronize()
    assert ctr.tolist() == [lost, esc]
    _same32(g, q)
    np.testing.assert_array_equal(dep.cpu().numpy(), dep_o)
    np.testing.assert_allclose(p_mac.cpu().numpy(), pm, rtol=1e-12, atol=1e-30)
    np.testing.assert_allclose(P_loss.item(), Pl, rtol=1e-12)


@pytest.mark.parametrize("layout", ["sorted", "unsorted"])
def test_deposit32_bitexact(gpu, oracle, rng, layout):
    """The fixed-point deposit of an FP32 set: bit-identical to the oracle for
    two particle orders (the unsorted set on a 40^3 grid overflows every
    block's LDS box), exactly N full weights on the grid; the FP64 density
    from it equals the oracle's."""
    from essaitemporaire_lepic_module_amd import particles32 as P32
    n, h = H.box((40, 40, 40) if layout == "unsorted" else (10, 9, 8))
    bc = cfg.grounded_box_bcnd(n)
    kq = cfg.build_kq(bc)
    count = 50000
    q, _ = _pair(oracle, P32, rng, n, h, count)
    if layout == "sorted":
        q = oracle.sort32(n, h, q)
    Nm = 1526.0
    S = _bits(n, h, Nm)
    ref = oracle.deposit32(n, h, q, Nm, S)
    a = q.arrays()
    for order in (np.arange(q.n), rng.permutation(q.n)):
        g = P32.ParticleSet32(q.n + 16, gpu).from_host(*(a[k][order] for k in oracle.FIELDS32))
        dep = torch.zeros(cfg.node_shape(n), dtype=torch.int64, device=gpu)
        P32.deposit32(g, n, h, Nm, S, dep)
        np.testing.assert_array_equal(dep.cpu().numpy(), ref)
    U = round(Nm / (h[0] * h[1] * h[2]) * 2.0 ** S)
    assert int(ref.sum()) == q.n * U
    npg = torch.zeros(cfg.node_shape(n), dtype=torch.float64, device=gpu)
    P32.fixed_density(n, h, dep, torch.as_tensor(kq, device=gpu), S, npg)
    np.testing.assert_array_equal(npg.cpu().numpy(), oracle.fixed_density(n, ref, kq, S))


@pytest.mark.parametrize("case", ["nearly_sorted", "unsorted_wide", "crowded_cell", "empty"])
def test_sort32_stable_bitexact(gpu, oracle, rng, case):
    """The device sort is the stable sort by cell with holes dropped: the
    oracle's order exactly, also for tiles spanning the whole grid (one
    atomic per particle) and for a cell holding more than 64 particles."""
    from essaitemporaire_lepic_module_amd import particles32 as P32
    n, h = H.box((40, 40, 40) if case == "unsorted_wide" else (9, 7, 6))
    count = {"nearly_sorted": 60000, "unsorted_wide": 60000, "crowded_cell": 20000,
             "empty": 0}[case]
    q, _ = _pair(oracle, P32, rng, n, h, count)
    a = q.arrays()
    if case == "nearly_sorted":
        q = oracle.sort32(n, h, q)
        a = q.arrays()
        # a few steps of drift: some particles one cell over
        shift = rng.random(q.n) < 0.2
        a["cell"][shift] = np.minimum(a["cell"][shift] + 1, n[0] * n[1] * n[2] - 1)
    if case == "crowded_cell":
        a["cell"][: 3000] = 17
    if count:
        a["cell"][rng.random(count) < 0.05] = -1
    q = oracle.Particles32(count + 16).set(*(a[k] for k in oracle.FIELDS32))
    g = P32.ParticleSet32(count + 16, gpu).from_host(*(a[k] for k in oracle.FIELDS32))
    ref = oracle.sort32(n, h, q)
    P32.sort32(g, n, h)
    assert g.n == ref.n
    _same32(g, ref)


def test_collisions32_deposit_exact(gpu, oracle, rng):
    """MCC on FP32 sets (cfg2's argon chemistry): products and kills adjust
    the fixed-point density exactly -- afterwards it equals a fresh deposit
    of the surviving particles bit for bit -- and the event rate matches the
    FP64 sets' on the same particles within binomial bounds."""
    from essaitemporaire_lepic_module_amd import particles as P
    from essaitemporaire_lepic_module_amd import particles32 as P32
    case = cfg.cfg2_pic_mcc(n=8, np_cell=40)
    n, h = case.n, case.h
    tables = P.ReactionTables(case.chemistry, case.Nm, gpu)
    nt = len(case.tracked)
    parts64 = []
    parts32 = []
    for s in range(nt):
        count = 30000
        arr = list(H.random_particles(rng, n, h, count, vscale=float(case.vt0[s]) * 3))
        ps = P.ParticleSet(2 * count, gpu).from_host(*arr)
        parts64.append(ps)
        g, out = P32.ParticleSet32.from_pset64(ps, n, h, cap=2 * count)
        assert out == 0
        parts32.append(g)
    bits = [_bits(n, h, case.Nm[s]) for s in range(nt)]
    shape = cfg.node_shape(n)
    dep = torch.zeros((nt,) + shape, dtype=torch.int64, device=gpu)
    for s in range(nt):
        P32.deposit32(parts32[s], n, h, case.Nm[s], bits[s], dep[s])
    np_red = torch.full((nt,) + shape, 1e15, dtype=torch.float64, device=gpu)
    Pc32 = torch.zeros(case.chemistry["ntype_total"], dtype=torch.float64, device=gpu)
    Pc64 = torch.zeros_like(Pc32)
    st32 = P32.collisions32(parts32, n, h, np_red, tables, 10, case.dt, Pc32, 0x1234, 7, dep,
                            bits)
    st64 = P.perform_collisions_step(parts64, n, h, np_red, tables, 10, case.dt, None, Pc64,
                                     mode=1, philox_key=0x1234, philox_counter=7)
    assert sum(st32["accepted"]) > 100
    for s in range(nt):
        a32, a64 = st32["accepted"][s], st64["accepted"][s]
        sd = math.sqrt(max(a64, 1))
        assert abs(a32 - a64) < 6 * sd + 5, (s, a32, a64)
        fresh = torch.zeros(shape, dtype=torch.int64, device=gpu)
        P32.deposit32(parts32[s], n, h, case.Nm[s], bits[s], fresh)
        assert torch.equal(fresh, dep[s])


class _OracleStep32:
    """Simulation32's step without collisions (sort at the reference's
    cadence plus the electron re-sort, density, the FP64 multigrid, E rounded
    to FP32, the FP32 push with the fixed-point deposit) driven through the
    oracle's routines on the host."""

    def __init__(self, oracle, case, parts, resort):
        self.O, self.c = oracle, case
        n, h = case.n, case.h
        self.bc = np.ascontiguousarray(case.bcnd, np.int32)
        self.kq = oracle.build_kq(n, self.bc)
        self.phi = case.phi0()
        self.tr = case.tracked
        self.Nm = case.Nm[: len(self.tr)]
        self.bits = [_bits(n, h, nm) for nm in self.Nm]
        self.resort = resort
        self.parts = []
        self.np_thread = np.zeros((len(self.tr),) + self.bc.shape)
        for s, a in enumerate(parts):
            p64 = oracle.Particles(len(a[0]) + 16).set(*a)
            q, out = oracle.from64(n, h, p64)
            assert out == 0
            q = oracle.sort32(n, h, q)
            self.parts.append(q)
            dep = oracle.deposit32(n, h, q, self.Nm[s], self.bits[s])
            self.np_thread[s] = oracle.fixed_density(n, dep, self.kq, self.bits[s])

    def step(self, istep):
        O, c = self.O, self.c
        n, h = c.n, c.h
        for s in range(len(self.parts)):
            if istep % c.nb_step_sort == 1 or (self.resort[s] and (istep - 1) % self.resort[s] == 0):
                self.parts[s] = O.sort32(n, h, self.parts[s])
        np_red = O.reduce_density(n, self.bc, self.np_thread)
        rho = O.build_rho(n, np_red, [sp.charge for sp in self.tr])
        st, _ = O.pdesolver(n, h, self.phi, rho, self.bc, eps=c.eps, omega=c.omega, ng=c.ng)
        assert st == 0
        E4f = O.round4(O.efield(n, h, self.phi, self.bc))
        ngrid = int(self.bc.max())
        for s, (sp, p) in enumerate(zip(self.tr, self.parts)):
            dep = np.zeros(self.bc.shape, np.int64)
            O.push32(n, h, p, E4f, 1, c.B_uniform, None, sp.charge, sp.mass, c.dt, self.bc, 0, 0,
                     s + 1, -1, self.Nm[s] * sp.charge, sp.mass, self.Nm[s], ngrid,
                     self.bits[s], dep)
            self.np_thread[s] = O.fixed_density(n, dep, self.kq, self.bits[s])


@pytest.mark.parametrize("poisson,tol", [("fp64", 1e-5), ("pcg64", 1e-3)])
def test_fp32_steps_match_oracle(gpu, oracle, rng, poisson, tol):
    """Simulation32 (the headline's step) without collisions, 25 steps of the
    8^3 magnetized box from the same particles (reference sorts at steps 1
    and 21, electrons re-sorted every 4 steps), against the oracle's FP32
    restatement step for step.  E is rounded to FP32 from potentials that
    agree to the solve's tolerance, so a value lying within that distance of
    a rounding boundary may round the other way: the particles are compared
    after a final sort, every one in the same cell and within tol of a cell
    (offsets) and tol relative (velocities), the same count alive, and the
    densities within tol of their maximum.  The conjugate-gradient solve
    (pcg64) stops on the same residual criterion from another iterate, so its
    fields, and the particles, differ by more."""
    from essaitemporaire_lepic_module_amd import particles as P
    from essaitemporaire_lepic_module_amd import particles32 as P32
    from essaitemporaire_lepic_module_amd.simulation import Simulation32
    case = cfg.cfg2_pic_mcc(n=8, np_cell=20, ng=3)
    case.chemistry = None
    n, h = case.n, case.h
    N = 20 * 8 ** 3
    parts = []
    for sp in case.tracked:
        vt = cfg.thermal_speed(sp.T_eV, sp.mass)
        parts.append(H.random_particles(rng, n, h, N, margin=0.01, vscale=vt / math.sqrt(2.0)))
    o = _OracleStep32(oracle, case, parts, [4, 0])
    p64 = [P.ParticleSet(2 * N, gpu).from_host(*a) for a in parts]
    sim = Simulation32(case, gpu, capacity_factor=2.0, resort_every=[4, 0], poisson=poisson,
                       parts64=p64)
    for k in range(1, 26):
        sim.advance_one_step(k)
        o.step(k)
    sim.check_escaped()
    np_g = sim.np_thread.cpu().numpy()
    for s in range(2):
        a = oracle.sort32(n, h, o.parts[s]).arrays()
        gq = oracle.Particles32(sim.parts[s].n + 16)
        gq.set(*(sim.parts[s].to_host()[k] for k in oracle.FIELDS32))
        b = oracle.sort32(n, h, gq).arrays()
        assert len(a["cell"]) == len(b["cell"])
        assert s == 1 or len(a["cell"]) < N          # the walls took electrons
        np.testing.assert_array_equal(a["cell"], b["cell"])
        for k in ("fx", "fy", "fz"):
            assert np.abs(a[k] - b[k]).max() <= tol, k
        for k in ("vx", "vy", "vz"):
            assert np.abs(a[k] - b[k]).max() <= tol * np.abs(a[k]).max(), k
        exact = np.mean(np.all([a[k] == b[k] for k in oracle.FIELDS32], axis=0))
        print(f"{poisson} species {s}: {len(a['cell'])} alive, {exact:.4f} bit-identical")
        assert np.abs(np_g[s] - o.np_thread[s]).max() <= tol * np.abs(o.np_thread[s]).max()
