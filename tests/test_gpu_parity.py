"""GPU parity: each liblepic_hip.so entry point against the oracle on the same
seeded inputs.  Integer outputs (cell ids, order, counts) and element-wise
FP64 results must be bit-identical; atomically accumulated fields
(deposition, wall diagnostics) agree to 1e-12 relative."""
import numpy as np
import pytest
import torch

from essaitemporaire_lepic_module_amd import config as cfg
from tests import helpers as H

pytestmark = pytest.mark.gpu

RTOL_ATOMIC = 1e-12


def _gpu_set(P, arrays, cap=None, dead=None):
    x, y, z, vx, vy, vz = arrays
    ps = P.ParticleSet(cap or len(x) + 16, "cuda:0")
    ps.from_host(x, y, z, vx, vy, vz, dead=dead)
    return ps


def _oracle_set(O, arrays, cap=None, dead=None):
    x = arrays[0]
    p = O.Particles(cap or len(x) + 16)
    p.set(*arrays, dead=dead)
    return p


def _same(a, b):
    assert a.shape == b.shape
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("n", [(8, 8, 8), (12, 10, 6)])
def test_move_electrostatic_bitexact(gpu, oracle, rng, n):
    from essaitemporaire_lepic_module_amd import particles as P
    n, h = H.box(n)
    arr = H.random_particles(rng, n, h, 5000, outside=0.02, vscale=3e5)
    E4 = H.random_E4(rng, n)
    ps = _gpu_set(P, arr)
    P.move_particles_electrostatic(ps, n, h, torch.as_tensor(E4, device=gpu), -1.6e-19, 9.1e-31, 5e-11)
    po = _oracle_set(oracle, arr)
    oracle.move_electrostatic(n, h, po, E4, -1.6e-19, 9.1e-31, 5e-11)
    g, o = ps.to_host(), po.arrays()
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        _same(g[k], o[k])


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_move_boris_bitexact(gpu, oracle, rng, mode):
    from essaitemporaire_lepic_module_amd import particles as P
    n, h = H.box((10, 8, 8))
    arr = H.random_particles(rng, n, h, 4000, vscale=1e6)
    E4 = H.random_E4(rng, n)
    B4 = np.zeros(cfg.node_shape(n) + (4,))
    B4[..., :3] = 0.02 * rng.standard_normal(cfg.node_shape(n) + (3,))
    B4[..., 3] = np.sqrt((B4[..., :3] ** 2).sum(-1))
    B4[..., 3][rng.random(B4.shape[:3]) < 0.3] = 0.0   # gate some nodes off
    Bc = (0.001, -0.002, 0.01)
    for q, m, magnetized in ((-1.6e-19, 9.1e-31, True), (1.6e-19, 6.6e-26, False)):
        ps = _gpu_set(P, arr)
        bf = P.bfield(mode, Bc, torch.as_tensor(B4, device=gpu) if mode == 2 else None, magnetized)
        P.move_particles_boris(ps, n, h, torch.as_tensor(E4, device=gpu), bf, q, m, 5e-11)
        po = _oracle_set(oracle, arr)
        oracle.move_boris(n, h, po, E4, mode, Bc, B4, q, m, 5e-11, magnetized)
        g, o = ps.to_host(), po.arrays()
        for k in ("x", "y", "z", "vx", "vy", "vz"):
            _same(g[k], o[k])


@pytest.mark.parametrize("nB", [(5, 4, 3), (10, 8, 8)])
def test_move_boris_bmap_own_grid_bitexact(gpu, oracle, rng, nB):
    """B map on its own grid (lepic_bfield mode 3, the reference's
    flag_gridB = 1), coarser than the field grid or of the same size."""
    from essaitemporaire_lepic_module_amd import particles as P
    n, h = H.box((10, 8, 8))
    hB = tuple(n[d] * h[d] / nB[d] for d in range(3))
    arr = H.random_particles(rng, n, h, 4000, vscale=1e6)
    E4 = H.random_E4(rng, n)
    shape = (nB[2] + 3, nB[1] + 3, nB[0] + 3)
    B4 = np.zeros(shape + (4,))
    B4[..., :3] = 0.02 * rng.standard_normal(shape + (3,))
    B4[..., 3] = np.sqrt((B4[..., :3] ** 2).sum(-1))
    B4[..., 3][rng.random(shape) < 0.3] = 0.0
    for q, m, magnetized in ((-1.6e-19, 9.1e-31, True), (1.6e-19, 6.6e-26, False)):
        ps = _gpu_set(P, arr)
        bf = P.bfield(3, B4=torch.as_tensor(B4, device=gpu), magnetized=magnetized, nB=nB, hB=hB)
        P.move_particles_boris(ps, n, h, torch.as_tensor(E4, device=gpu), bf, q, m, 5e-11)
        po = _oracle_set(oracle, arr)
        oracle.move_boris_bgrid(n, h, po, E4, B4, nB, hB, q, m, 5e-11, magnetized)
        g, o = ps.to_host(), po.arrays()
        for k in ("x", "y", "z", "vx", "vy", "vz"):
            _same(g[k], o[k])


def test_leapfrog_equals_electrostatic(gpu, rng):
    """Boris with B off is the modular leapfrog bit for bit (2*k1 == dt*q/m)."""
    from essaitemporaire_lepic_module_amd import particles as P
    n, h = H.box((8, 8, 8))
    arr = H.random_particles(rng, n, h, 3000)
    E4 = torch.as_tensor(H.random_E4(rng, n), device=gpu)
    a, b = _gpu_set(P, arr), _gpu_set(P, arr)
    P.move_particles_electrostatic(a, n, h, E4, -1.6e-19, 9.1e-31, 5e-11)
    P.move_particles_boris(b, n, h, E4, P.bfield(0), -1.6e-19, 9.1e-31, 5e-11)
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        _same(a.to_host()[k], b.to_host()[k])


@pytest.mark.parametrize("bmaker,pbc", [(cfg.grounded_box_bcnd, 0), (H.two_electrode_bcnd, 0),
                                        (H.periodic_yz_bcnd, 1)])
def test_apply_bc_bitexact(gpu, oracle, rng, bmaker, pbc):
    from essaitemporaire_lepic_module_amd import particles as P
    n, h = H.box((8, 8, 8))
    bc = bmaker(n)
    ngrid = int(bc.max())
    count = 20000
    arr = H.random_particles(rng, n, h, count, outside=0.05)
    dead = (rng.random(count) < 0.05).astype(np.int8)
    ps = _gpu_set(P, arr, dead=dead)
    p_mac = torch.zeros((ngrid + 1, 2), dtype=torch.float64, device=gpu)
    P_loss = torch.zeros(1, dtype=torch.float64, device=gpu)
    lost = P.apply_particle_bc_legacy(ps, n, h, torch.as_tensor(bc, device=gpu), pbc, 0, 1, -1, 0,
                                      -2.4e-16, p_mac, 9.1e-31, P_loss, 1500.0)
    po = _oracle_set(oracle, arr, dead=dead)
    lost_o, pmac_o, Ploss_o = oracle.apply_bc(n, h, po, bc, pbc, 0, 1, -1, -2.4e-16, 9.1e-31,
                                              1500.0, ngrid)
    assert lost == lost_o
    assert ps.n == po.n
    g, o = ps.to_host(), po.arrays()
    for k in ("x", "y", "z", "vx", "vy", "vz", "dead"):
        _same(g[k], o[k])
    np.testing.assert_allclose(p_mac.cpu().numpy().reshape(-1), pmac_o, rtol=RTOL_ATOMIC, atol=0)
    np.testing.assert_allclose(P_loss.item(), Ploss_o, rtol=RTOL_ATOMIC)


@pytest.mark.parametrize("count", [0, 1, 255, 256, 257, 70000])
def test_apply_bc_ragged_sizes(gpu, oracle, rng, count):
    """Tile edges of the look-back compaction: empty, one, partial tiles."""
    from essaitemporaire_lepic_module_amd import particles as P
    n, h = H.box((6, 6, 6))
    bc = cfg.grounded_box_bcnd(n)
    arr = H.random_particles(rng, n, h, count, outside=0.1)
    ps = _gpu_set(P, arr)
    p_mac = torch.zeros((2, 2), dtype=torch.float64, device=gpu)
    P_loss = torch.zeros(1, dtype=torch.float64, device=gpu)
    lost = P.apply_particle_bc_legacy(ps, n, h, torch.as_tensor(bc, device=gpu), 0, 0, 1, -1, 0,
                                      1.0, p_mac, 1.0, P_loss, 1.0)
    po = _oracle_set(oracle, arr)
    lost_o, _, _ = oracle.apply_bc(n, h, po, bc, 0, 0, 1, -1, 1.0, 1.0, 1.0, 1)
    assert lost == lost_o and ps.n == po.n
    g, o = ps.to_host(), po.arrays()
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        _same(g[k], o[k])


def test_sort_bitexact(gpu, oracle, rng):
    from essaitemporaire_lepic_module_amd import particles as P
    n, h = H.box((16, 8, 12))
    arr = H.random_particles(rng, n, h, 50000, outside=0.01)
    dead = (rng.random(50000) < 0.1).astype(np.int8)
    ps = _gpu_set(P, arr, dead=dead)
    P.sort_particles_by_cell(ps, n, h)
    po = _oracle_set(oracle, arr, dead=dead)
    cc, cs = oracle.sort_particles(n, h, po)
    g, o = ps.to_host(), po.arrays()
    for k in ("x", "y", "z", "vx", "vy", "vz", "cell", "dead"):
        _same(g[k], o[k])
    _same(ps.cell_count.cpu().numpy(), cc)
    _same(ps.cell_start.cpu().numpy(), cs)


def test_deposit(gpu, oracle, rng):
    from essaitemporaire_lepic_module_amd import particles as P, fields as F
    n, h = H.box((8, 8, 8))
    bc = cfg.grounded_box_bcnd(n)
    kq = cfg.build_kq(bc)
    for sort_first in (False, True):
        arr = H.random_particles(rng, n, h, 40000, outside=0.01)
        dead = (rng.random(40000) < 0.05).astype(np.int8)
        ps = _gpu_set(P, arr, dead=dead)
        if sort_first:   # sorted tiles exercise the LDS box path
            P.sort_particles_by_cell(ps, n, h)
        npg = torch.zeros(cfg.node_shape(n), dtype=torch.float64, device=gpu)
        F.deposit_particle_set_to_np_thread(ps, n, h, torch.as_tensor(kq, device=gpu), 1526.0, npg)
        po = _oracle_set(oracle, arr, dead=dead)
        npo = oracle.deposit(n, h, po, kq, 1526.0)
        np.testing.assert_allclose(npg.cpu().numpy(), npo, rtol=RTOL_ATOMIC, atol=1e-30)


def test_kq_reduce_rho_bitexact(gpu, oracle, rng):
    from essaitemporaire_lepic_module_amd import fields as F
    n, h = H.box((8, 6, 10))
    for bc in (cfg.grounded_box_bcnd(n), H.periodic_yz_bcnd(n)):
        kq_o = oracle.build_kq(n, bc)
        kq_g = F.build_kq(torch.as_tensor(bc, device=gpu), n)
        _same(kq_g.cpu().numpy(), kq_o)
        npt = rng.random((2,) + cfg.node_shape(n)) * 1e15
        npt_g = torch.as_tensor(npt.copy(), device=gpu)
        np_red_g = torch.zeros_like(npt_g)
        F.reduce_species_density(n, torch.as_tensor(bc, device=gpu), npt_g, 2, np_red_g)
        np_red_o = oracle.reduce_density(n, bc, npt)
        _same(np_red_g.cpu().numpy(), np_red_o)
        _same(npt_g.cpu().numpy(), npt)   # stitching is in place in both
        q = [-1.60217646e-19, 1.60217646e-19]
        rho_g = torch.zeros(cfg.node_shape(n), dtype=torch.float64, device=gpu)
        F.build_rho_from_np(n, np_red_g, q, 2, rho_g)
        _same(rho_g.cpu().numpy(), oracle.build_rho(n, np_red_o, q))


@pytest.mark.parametrize("bmaker", [cfg.grounded_box_bcnd, H.two_electrode_bcnd, H.periodic_yz_bcnd,
                                    H.thin_gap_bcnd])
def test_efield_bitexact(gpu, oracle, rng, bmaker):
    from essaitemporaire_lepic_module_amd import fields as F
    n, h = H.box((8, 8, 8))
    bc = bmaker(n)
    phi = rng.standard_normal(cfg.node_shape(n))
    E_g = torch.zeros(cfg.node_shape(n) + (4,), dtype=torch.float64, device=gpu)
    F.calc_Efield_modular(n, h, torch.as_tensor(phi, device=gpu), E_g, torch.as_tensor(bc, device=gpu))
    _same(E_g.cpu().numpy(), oracle.efield(n, h, phi, bc))


# The solver runs the levels below ~12000 nodes inside one workgroup: the 16^3
# cases put every level but the finest there, 64^3 and 128^3 keep two and
# three levels on the multi-block path, ng = 1 has no coarse part at all.
@pytest.mark.parametrize("bmaker,pbc,ng,size", [(cfg.grounded_box_bcnd, 0, 4, 16),
                                                (H.two_electrode_bcnd, 0, 3, 16),
                                                (H.periodic_yz_bcnd, 1, 3, 16),
                                                (cfg.grounded_box_bcnd, 0, 1, 8),
                                                (H.periodic_yz_bcnd, 1, 4, 32),
                                                (H.two_electrode_bcnd, 0, 5, 64),
                                                (cfg.grounded_box_bcnd, 0, 7, 128)])
def test_poisson_bitexact(gpu, oracle, rng, bmaker, pbc, ng, size):
    from essaitemporaire_lepic_module_amd import fields as F
    n, h = H.box((size, size, size))
    bc = bmaker(n)
    rho = np.zeros(cfg.node_shape(n))
    rho[bc <= 0] = 1e-4 * rng.standard_normal(int((bc <= 0).sum()))
    phi0 = np.zeros(cfg.node_shape(n))
    phi0[bc == 2] = 20.0
    phi0[bc == 3] = -5.0
    phi_g = torch.as_tensor(phi0.copy(), device=gpu)
    info_g = F.solve_poisson_legacy(phi_g, torch.as_tensor(bc, device=gpu),
                                    torch.as_tensor(rho, device=gpu), h, n, 20000, 1e-8, 1.4, ng,
                                    pbc, 0)
    phi_o = phi0.copy()
    st, info_o = oracle.pdesolver(n, h, phi_o, rho, bc, flag_pbc=pbc, ng=ng, eps=1e-8, omega=1.4)
    assert st == 0
    assert info_g["cycles"] == info_o["cycles"]
    assert info_g["sweeps"] == info_o["sweeps"]
    np.testing.assert_allclose(phi_g.cpu().numpy(), phi_o, rtol=1e-12, atol=1e-12 * np.abs(phi_o).max())
    np.testing.assert_allclose(info_g["res"], info_o["res"], rtol=1e-9)


def test_loader_host_bitexact(oracle):
    """The product's host loader reproduces the reference loader's ran2 stream."""
    from essaitemporaire_lepic_module_amd import particles as P
    import torch as _t
    if not _t.cuda.is_available():
        pytest.skip("needs the device")
    c = cfg.cfg1_plain_box(n=8, np_cell=5)
    iseed = [123456]
    parts, npt = P.load_particles_modular(c.n, c.h, c.bcnd, c.kq(), c.np_cell, c.vt0[:2], c.Nm[:2],
                                          [1.0, 1.0], iseed, cap=4000)
    po, npo, so = oracle.load_particles(c.n, c.h, c.bcnd, c.kq(), c.np_cell, c.vt0[:2], c.Nm[:2],
                                        [1.0, 1.0], 123456, cap=4000)
    assert iseed[0] == so
    for a, b in zip(parts, po):
        g, o = a.to_host(), b.arrays()
        assert len(g["x"]) == len(o["x"])
        for k in ("x", "y", "z"):
            _same(g[k], o[k])
        # velocities pass through log/cos/sin: the host libm may differ by an ulp
        for k in ("vx", "vy", "vz"):
            np.testing.assert_allclose(g[k], o[k], rtol=1e-15, atol=0)
    _same(npt.cpu().numpy(), npo)


@pytest.mark.parametrize("neg", [False, True])
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("presort,size", [(True, (12, 10, 8)), (False, (12, 10, 8)),
                                          (False, (20, 20, 20))])
def test_fused_inplace_and_sort_compact(gpu, oracle, rng, mode, presort, size, neg):
    """Production particle phase: in-place push + wall BC + deposit with holes,
    then the compacting sort, against the oracle's mover + BC + deposit + sort.
    Sorted input deposits through the LDS box in runs of one cell; unsorted
    input has no runs, and on the 20^3 grid a block's particles span more
    nodes than the box holds, so every stencil goes straight to global memory.
    neg makes the particles negative ions with a Neumann x = 0 face: those
    that cross it are lost (lost = 2, no wall label)."""
    from essaitemporaire_lepic_module_amd import particles as P
    n, h = H.box(size)
    bc = H.two_electrode_bcnd(n)
    if neg:   # a Neumann x = 0 face (-2): the cells next to it are not walls
        bc[2:n[2] + 1, 2:n[1] + 1, 0:2] = -2
    kq = cfg.build_kq(bc)
    count = 30000
    arr = H.random_particles(rng, n, h, count, outside=0.02, vscale=2e6)
    dead = (rng.random(count) < 0.03).astype(np.int8)
    E4 = H.random_E4(rng, n)
    q, m, dt, Nm = -1.60217646e-19, 9.10938188e-31, 5e-11, 1526.0
    Bc = (0.0, 0.0, 0.01)
    po = _oracle_set(oracle, arr, dead=dead)
    if presort:
        oracle.sort_particles(n, h, po)
    start = po.arrays()
    ps = _gpu_set(P, [start[k] for k in ("x", "y", "z", "vx", "vy", "vz")], dead=start["dead"])
    bf = P.bfield(mode, Bc)
    nmn, ptype, tag_neg = (1, 2, 2) if neg else (0, 1, -1)
    bp = P.bc_params(n, h, 0, nmn, ptype, tag_neg, Nm * q, m, Nm, int(bc.max()))
    p_mac = torch.zeros((int(bc.max()) + 1, 2), dtype=torch.float64, device=gpu)
    P_loss = torch.zeros(1, dtype=torch.float64, device=gpu)
    npg = torch.zeros(cfg.node_shape(n), dtype=torch.float64, device=gpu)
    bcg = torch.as_tensor(bc, device=gpu)
    lost = P.push_bc_deposit_inplace(ps, n, h, torch.as_tensor(E4, device=gpu), bf, q, m, dt, bcg,
                                     bp, torch.as_tensor(kq, device=gpu), npg, p_mac, P_loss)
    # oracle: mover, BC (removes MCC-dead too), deposit, sort
    if mode == 0:
        oracle.move_electrostatic(n, h, po, E4, q, m, dt)
    else:
        oracle.move_boris(n, h, po, E4, mode, Bc, None, q, m, dt)
    lost_o, pmac_o, Ploss_o = oracle.apply_bc(n, h, po, bc, 0, nmn, ptype, tag_neg, Nm * q, m, Nm,
                                              int(bc.max()))
    npo = oracle.deposit(n, h, po, kq, Nm)
    assert lost == lost_o - int(start["dead"].sum())   # MCC-dead holes are not new losses
    g = ps.to_host()
    keep = g["dead"] == 0
    o = po.arrays()
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        _same(g[k][keep], o[k])
    np.testing.assert_allclose(npg.cpu().numpy(), npo, rtol=RTOL_ATOMIC, atol=1e-30)
    np.testing.assert_allclose(p_mac.cpu().numpy().reshape(-1), pmac_o, rtol=RTOL_ATOMIC, atol=0)
    P.sort_compact(ps, n, h)
    oracle.sort_particles(n, h, po)
    g, o = ps.to_host(), po.arrays()
    assert ps.n == po.n
    for k in ("x", "y", "z", "vx", "vy", "vz", "cell"):
        _same(g[k], o[k])


@pytest.mark.parametrize("presort", [False, True])
def test_deposit_fixed_bitexact(gpu, oracle, rng, presort):
    """The fixed-point deposit is bit-identical to the oracle's (integer sums
    do not depend on order), also for the particles in another order, and
    rejects a scale outside [1, 2^48] units per full weight."""
    from essaitemporaire_lepic_module_amd import fields as F
    from essaitemporaire_lepic_module_amd import particles as P
    from essaitemporaire_lepic_module_amd._native import LepicError
    n, h = H.box((12, 10, 8))
    bc = H.two_electrode_bcnd(n)
    kq = cfg.build_kq(bc)
    count = 30000
    arr = H.random_particles(rng, n, h, count, outside=0.02)
    dead = (rng.random(count) < 0.03).astype(np.int8)
    po = _oracle_set(oracle, arr, dead=dead)
    if presort:
        oracle.sort_particles(n, h, po)
    a = po.arrays()
    cols = [a[k] for k in ("x", "y", "z", "vx", "vy", "vz")]
    Nm = 1526.0
    kqg = torch.as_tensor(kq, device=gpu)
    outs = []
    for order in (np.arange(po.n), rng.permutation(po.n)):
        ps = _gpu_set(P, [c[order] for c in cols], dead=a["dead"][order])
        npg = torch.zeros(cfg.node_shape(n), dtype=torch.float64, device=gpu)
        S = F.deposit_particle_set_to_np_thread_fixed(ps, n, h, kqg, Nm, npg)
        outs.append(npg.cpu().numpy())
    ref = oracle.deposit_fixed(n, h, po, kq, Nm, S)
    assert np.abs(ref).sum() > 0
    _same(outs[0], ref)
    _same(outs[1], ref)
    with pytest.raises(LepicError):
        F.deposit_particle_set_to_np_thread_fixed(ps, n, h, kqg, Nm, npg, frac_bits=S + 10)
    before = npg.clone()
    F.deposit_particle_set_to_np_thread_fixed(ps, n, h, kqg, 0.0, npg)   # zero weight: no-op
    assert torch.equal(npg, before)


def test_deposit_unsorted_falls_back(gpu, oracle, rng):
    """Particles spread over the whole grid in every block: the LDS box does
    not fit and every particle goes straight to the replicas."""
    from essaitemporaire_lepic_module_amd import particles as P, fields as F
    n, h = H.box((40, 40, 40))
    bc = cfg.grounded_box_bcnd(n)
    kq = cfg.build_kq(bc)
    arr = H.random_particles(rng, n, h, 20000)
    ps = _gpu_set(P, arr)
    npg = torch.zeros(cfg.node_shape(n), dtype=torch.float64, device=gpu)
    F.deposit_particle_set_to_np_thread(ps, n, h, torch.as_tensor(kq, device=gpu), 3.0, npg)
    npo = oracle.deposit(n, h, _oracle_set(oracle, arr), kq, 3.0)
    np.testing.assert_allclose(npg.cpu().numpy(), npo, rtol=RTOL_ATOMIC, atol=1e-30)


@pytest.mark.parametrize("layout", ["unsorted", "sorted_far"])
def test_deposit_fixed_replica_path_bitexact(gpu, oracle, rng, layout):
    """The fixed-point deposit's paths that bypass the LDS box: unsorted
    particles on a 40^3 grid (every block's box overflows, so every stencil
    goes to the int64 replicas), and a sorted set salted with far-away
    particles (their stencils leave the box of a block whose box fits).
    Bit-identical to the oracle in two particle orders, and the charge on the
    grid is exactly the deposited count times one full weight."""
    from essaitemporaire_lepic_module_amd import fields as F
    from essaitemporaire_lepic_module_amd import particles as P
    n, h = H.box((40, 40, 40))
    bc = cfg.grounded_box_bcnd(n)
    kq = cfg.build_kq(bc)
    count = 60000
    arr = [a.copy() for a in H.random_particles(rng, n, h, count)]
    po = _oracle_set(oracle, arr)
    if layout == "sorted_far":
        oracle.sort_particles(n, h, po)
        a = po.arrays()
        arr = [a[k].copy() for k in ("x", "y", "z", "vx", "vy", "vz")]
        far = rng.choice(count, 300, replace=False)
        L = np.array(n) * np.array(h)
        for d in range(3):
            arr[d][far] = L[d] * rng.random(far.size)
        po = _oracle_set(oracle, arr)
    Nm = 1526.0
    kqg = torch.as_tensor(kq, device=gpu)
    S = None
    for order in (np.arange(count), rng.permutation(count)):
        ps = _gpu_set(P, [c[order] for c in arr])
        npg = torch.zeros(cfg.node_shape(n), dtype=torch.float64, device=gpu)
        S = F.deposit_particle_set_to_np_thread_fixed(ps, n, h, kqg, Nm, npg)
        ref = oracle.deposit_fixed(n, h, po, kq, Nm, S)
        _same(npg.cpu().numpy(), ref)
    k1 = Nm / (h[0] * h[1] * h[2])
    U = round(k1 * 2.0 ** S)
    units = np.rint(ref / kq * 2.0 ** S).astype(np.int64)
    assert int(units.sum()) == count * U


@pytest.mark.parametrize("case", ["mixed", "empty", "all_dead", "one_cell", "wide"])
def test_bin_compact_groups_by_cell(gpu, oracle, rng, case):
    """Binning keeps exactly the live particles, in nondecreasing cell order,
    each cell holding the same particles as the stable sort puts there."""
    from essaitemporaire_lepic_module_amd import particles as P
    # "wide": unsorted input on a larger grid, so that every tile's cell box
    # exceeds the staged path and takes one atomic per particle
    n, h = H.box((40, 40, 40) if case == "wide" else (9, 7, 6))
    count = {"mixed": 40000, "empty": 0, "all_dead": 3000, "one_cell": 5000, "wide": 60000}[case]
    arr = H.random_particles(rng, n, h, count, outside=0.02)
    if case == "one_cell":
        arr = [a.copy() for a in arr]
        for d in range(3):
            arr[d][:] = 2.5 * h[d] + 0.1 * h[d] * rng.random(count)
    dead = (rng.random(count) < 0.05).astype(np.int8) * 2
    if case == "all_dead":
        dead[:] = 2
    # nearly sorted input, as in production
    po = _oracle_set(oracle, arr, dead=np.zeros(count, np.int8))
    if count and case != "wide":
        oracle.sort_particles(n, h, po)
    start = po.arrays()
    a = _gpu_set(P, [start[k] for k in ("x", "y", "z", "vx", "vy", "vz")], dead=dead)
    b = _gpu_set(P, [start[k] for k in ("x", "y", "z", "vx", "vy", "vz")], dead=dead)
    P.bin_compact(a, n, h)
    P.sort_compact(b, n, h)
    assert a.n == b.n == int((dead == 0).sum())
    ga, gb = a.to_host(), b.to_host()
    assert (ga["dead"] == 0).all()
    assert (np.diff(ga["cell"]) >= 0).all()
    np.testing.assert_array_equal(ga["cell"], gb["cell"])
    cols = ("cell", "x", "y", "z", "vx", "vy", "vz")
    ka = np.stack([ga[k].astype(np.float64) for k in cols], 1)
    kb = np.stack([gb[k].astype(np.float64) for k in cols], 1)
    np.testing.assert_array_equal(ka[np.lexsort(ka.T[::-1])], kb[np.lexsort(kb.T[::-1])])


@pytest.mark.parametrize("bmaker,pbc,ng,size", [(cfg.grounded_box_bcnd, 0, 5, 64),
                                                (H.periodic_yz_bcnd, 1, 4, 32)])
def test_poisson_graph_replay_matches_eager(gpu, rng, bmaker, pbc, ng, size):
    """V-cycles replayed from the captured HIP graph give the same iterates as
    the same launches issued one by one: several warm-started solves in a row
    (each replaying the graph several times), eager work and host syncs
    between them, and a second right-hand side buffer that forces a new
    capture."""
    from essaitemporaire_lepic_module_amd import fields as F
    n, h = H.box((size, size, size))
    bc = torch.as_tensor(bmaker(n), device=gpu)
    phis = {g: torch.zeros(cfg.node_shape(n), dtype=torch.float64, device=gpu) for g in (True, False)}
    rhos = [torch.zeros(cfg.node_shape(n), dtype=torch.float64, device=gpu) for _ in range(2)]
    interior = bc <= 0
    for it in range(6):
        rho = rhos[it // 3]
        rho.zero_()
        rho[interior] = 1e-4 * torch.as_tensor(rng.standard_normal(int(interior.sum())), device=gpu)
        info = {}
        for graph in (True, False):
            info[graph] = F.solve_poisson_legacy(phis[graph], bc, rho, h, n, 20000, 1e-8, 1.4, ng,
                                                 pbc, 0, graph=graph)
            torch.cuda.synchronize()
        assert info[True] == info[False], it
        assert info[True]["cycles"] >= 2
        assert torch.equal(phis[True], phis[False]), it
