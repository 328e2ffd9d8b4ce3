"""GPU null-collision MCC (perform_collisions_step, mod_collisions.f90:214-739).

REPLAY_RAN2 runs the reference's sequential chain and must agree with the
oracle on the event stream exactly: counters, the ran2 seed after the chain,
product slots, positions and flags.  Post-collision velocities and Pcoll go
through log, sin and cos (box_muller, scatter_vector), which neither the
host libm nor the GPU's math library rounds correctly, so they are compared
to 1e-12 relative (the two libraries differ in the last bit).  PHILOX is
the parallel production mode: it is checked against the laws of the
reference's process (expected number of accepted collisions, dead particles
untouched, speed kept by elastic collisions on a cold heavy target, one e-
and one Ar+ per ionization, products on their parent's position) and for
determinism under a fixed (key, counter)."""
import math

import numpy as np
import pytest
import torch

from essaitemporaire_lepic_module_amd import _native
from essaitemporaire_lepic_module_amd import config as cfg
from tests import helpers as H

pytestmark = pytest.mark.gpu

QE = 1.60217646e-19
RTOL_TRANSCENDENTAL = 1e-12


def _elastic_tables(ngas, sigma, T_eV=1e-6):
    sp = [cfg.Species("e", cfg.ME, -QE, 1.0),
          cfg.Species("Ar", -39.95 * cfg.AMU, 0.0, T_eV, tracked=False)]
    E = cfg.argon_energy_grid(120)
    return cfg.build_reaction_tables(sp, ngas, [((0, 1), (0, 1), 0.0, None)],
                                     [(E, np.full_like(E, sigma))])


def _pair(P, O, arrays, cap, dead=None):
    ps = P.ParticleSet(cap, "cuda:0")
    ps.from_host(*arrays, dead=dead)
    po = O.Particles(cap).set(*arrays, dead=dead)
    return ps, po


def _run_gpu(P, parts, n, h, np_red, tables, ns, dt, mode, seed=123456, key=0, counter=0):
    ntot = tables.t["ntype_total"]
    Pcoll = torch.zeros(ntot, dtype=torch.float64, device="cuda:0")
    iseed = [seed]
    st = P.perform_collisions_step(parts, n, h, np_red, tables, ns, dt, iseed, Pcoll, mode=mode,
                                   philox_key=key, philox_counter=counter)
    return st, iseed[0], Pcoll.cpu().numpy()


def test_mcc_replay_bitexact(gpu, oracle, rng):
    from essaitemporaire_lepic_module_amd import particles as P
    c = cfg.cfg2_pic_mcc(n=4, np_cell=1)
    n, h = c.n, c.h
    t = dict(c.chemistry)
    t["Nm"] = np.array(c.Nm)
    N = 6000
    L = n[0] * h[0]
    # thermal electrons plus a 100-300 eV tail so that every channel fires
    ve = 1e6 * rng.standard_normal((N, 3))
    tail = rng.random(N) < 0.3
    speed = np.sqrt(2 * rng.uniform(100, 300, N) * QE / cfg.ME)
    dirs = rng.standard_normal((N, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    ve[tail] = (speed[:, None] * dirs)[tail]
    pos_e = L * rng.random((N, 3))
    pos_i = L * rng.random((N, 3))
    vi = 3e4 * rng.standard_normal((N, 3))      # fast enough for a few charge exchanges
    dead_e = (rng.random(N) < 0.05).astype(np.int8)
    cap = 4 * N
    e_g, e_o = _pair(P, oracle, [pos_e[:, 0], pos_e[:, 1], pos_e[:, 2], ve[:, 0], ve[:, 1], ve[:, 2]],
                     cap, dead_e)
    i_g, i_o = _pair(P, oracle, [pos_i[:, 0], pos_i[:, 1], pos_i[:, 2], vi[:, 0], vi[:, 1], vi[:, 2]],
                     cap)
    np_red = np.abs(rng.standard_normal((2,) + cfg.node_shape(n))) * 1e15
    st_o, seed_o, Pcoll_o, s_o = oracle.collisions(n, h, [e_o, i_o], np_red, oracle.ChemTables(t),
                                                   c.nb_step_collisions, c.dt, 123456)
    assert st_o == 0
    tables = P.ReactionTables(c.chemistry, c.Nm, gpu)
    s_g, seed_g, Pcoll_g = _run_gpu(P, [e_g, i_g], n, h, torch.as_tensor(np_red, device=gpu), tables,
                                    c.nb_step_collisions, c.dt, _native.MCC_REPLAY_RAN2)
    assert s_o["appended"][0] > 0 and s_o["accepted"][1] > 0
    for k in ("Nc", "accepted", "err_coll", "skipped_dead", "appended"):
        assert s_g[k] == s_o[k], k
    assert seed_g == seed_o
    np.testing.assert_allclose(Pcoll_g, Pcoll_o, rtol=RTOL_TRANSCENDENTAL, atol=0)
    for pg, po in ((e_g, e_o), (i_g, i_o)):
        assert pg.n == po.n
        g, o = pg.to_host(), po.arrays()
        for k in ("x", "y", "z", "dead"):
            np.testing.assert_array_equal(g[k], o[k], err_msg=k)
        vg = np.stack([g["vx"], g["vy"], g["vz"]], 1)
        vo = np.stack([o["vx"], o["vy"], o["vz"]], 1)
        scale = np.linalg.norm(vo, axis=1, keepdims=True)
        assert np.all(np.abs(vg - vo) <= RTOL_TRANSCENDENTAL * scale)


def _philox_elastic_setup(P, rng, N, gpu, dead_every=37):
    n, h = H.box((6, 6, 6))
    ngas, sigma, dt, ns = 2.4e20, 1e-19, 5e-11, 10
    t = _elastic_tables(ngas, sigma)
    L = n[0] * h[0]
    speed = 2e5 + 1.5e6 * rng.random(N)
    dirs = rng.standard_normal((N, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    v = speed[:, None] * dirs
    pos = L * rng.random((N, 3))
    dead = np.zeros(N, np.int8)
    dead[::dead_every] = 2          # holes left by the wall BC
    arrays = [pos[:, 0], pos[:, 1], pos[:, 2], v[:, 0], v[:, 1], v[:, 2]]
    tables = P.ReactionTables(t, [1.0, 1.0], gpu)
    np_red = torch.zeros((1,) + cfg.node_shape(n), dtype=torch.float64, device=gpu)
    return n, h, t, ngas, sigma, dt, ns, arrays, dead, speed, tables, np_red


def test_mcc_philox_elastic_rate_and_invariants(gpu, rng):
    from essaitemporaire_lepic_module_amd import particles as P
    N = 400_000
    n, h, t, ngas, sigma, dt, ns, arrays, dead, speed, tables, np_red = \
        _philox_elastic_setup(P, rng, N, gpu)
    ps = P.ParticleSet(N + 1024, "cuda:0")
    ps.from_host(*arrays, dead=dead)
    st, _, Pcoll = _run_gpu(P, [ps], n, h, np_red, tables, ns, dt, _native.MCC_PHILOX, key=77, counter=3)
    live = dead == 0
    # expected accepted collisions: each live particle takes Poisson(Pmax) trials, each accepted
    # with probability nu(v)/nu_max_local, nu_max_local = Pmax/(ns dt): sum_i ns dt ngas sigma v_i
    expect = float(np.sum(ns * dt * ngas * sigma * speed[live]))
    assert abs(st["accepted"][0] - expect) < 5 * math.sqrt(expect) + 5
    assert st["skipped_dead"][0] == 0 and st["appended"][0] == 0
    g = ps.to_host()
    assert ps.n == N
    vg = np.stack([g["vx"], g["vy"], g["vz"]], 1)
    v0 = np.stack(arrays[3:], 1)
    # holes are never touched
    np.testing.assert_array_equal(vg[~live], v0[~live])
    np.testing.assert_array_equal(g["dead"], dead)
    for k, a in zip(("x", "y", "z"), arrays[:3]):
        np.testing.assert_array_equal(g[k], a)
    # elastic on a cold argon atom keeps the electron's speed to ~2 me/M
    sp = np.linalg.norm(vg, axis=1)
    np.testing.assert_allclose(sp[live], speed[live], rtol=1e-4)
    changed = np.any(vg != v0, axis=1).sum()
    assert 0 < changed <= st["accepted"][0]
    # energy bookkeeping: Pcoll is the sum of 0.5 Nm m (v'^2 - v^2)
    de = 0.5 * cfg.ME * float(np.sum(sp[live] ** 2 - speed[live] ** 2))
    assert Pcoll[0] == pytest.approx(de, rel=1e-6, abs=1e-30)


def test_mcc_philox_deterministic(gpu, rng):
    from essaitemporaire_lepic_module_amd import particles as P
    N = 50_000
    n, h, t, ngas, sigma, dt, ns, arrays, dead, speed, tables, np_red = \
        _philox_elastic_setup(P, rng, N, gpu)
    out = []
    for counter in (5, 5, 6):
        ps = P.ParticleSet(N + 1024, "cuda:0")
        ps.from_host(*arrays, dead=dead)
        _run_gpu(P, [ps], n, h, np_red, tables, ns, dt, _native.MCC_PHILOX, key=9, counter=counter)
        out.append(ps.to_host())
    for k in ("vx", "vy", "vz"):
        np.testing.assert_array_equal(out[0][k], out[1][k])
    assert not np.array_equal(out[0]["vx"], out[2]["vx"])


def test_mcc_philox_ionization_products(gpu, rng):
    from essaitemporaire_lepic_module_amd import particles as P
    c = cfg.cfg2_pic_mcc(n=4, np_cell=1)
    n, h = c.n, c.h
    N = 100_000
    L = n[0] * h[0]
    speed = np.sqrt(2 * 200 * QE / cfg.ME)
    pos = L * rng.random((N, 3))
    e = P.ParticleSet(3 * N, "cuda:0")
    e.from_host(pos[:, 0], pos[:, 1], pos[:, 2], speed * np.ones(N), np.zeros(N), np.zeros(N))
    ipos = L * rng.random((N, 3))
    ion = P.ParticleSet(3 * N, "cuda:0")
    ion.from_host(ipos[:, 0], ipos[:, 1], ipos[:, 2], np.zeros(N), np.zeros(N), np.zeros(N))
    tables = P.ReactionTables(c.chemistry, c.Nm, gpu)
    np_red = torch.zeros((2,) + cfg.node_shape(n), dtype=torch.float64, device=gpu)
    st, _, _ = _run_gpu(P, [e, ion], n, h, np_red, tables, c.nb_step_collisions, c.dt,
                        _native.MCC_PHILOX, key=3, counter=1)
    a = st["appended"]
    assert a[0] == a[1] > 0
    assert e.n == N + a[0] and ion.n == N + a[1]
    ge, gi = e.to_host(), ion.to_host()
    parent = set(zip(ge["x"][:N], ge["y"][:N], ge["z"][:N]))
    prod_e = list(zip(ge["x"][N:], ge["y"][N:], ge["z"][N:]))
    prod_i = list(zip(gi["x"][N:], gi["y"][N:], gi["z"][N:]))
    assert all(p in parent for p in prod_e)
    assert sorted(prod_e) == sorted(prod_i)
    assert (ge["dead"][N:] == 0).all() and (gi["dead"][N:] == 0).all()


def test_mcc_philox_fast_path_matches_general_trial(gpu, rng):
    """Projectiles whose targets are all neutral take the register-only trial;
    adding a zero cross-section reaction on a charged target sends electrons
    through the general trial instead.  That reaction draws nothing and never
    fires, so both see the same draws and must make the same collisions."""
    from essaitemporaire_lepic_module_amd import particles as P
    c = cfg.cfg2_pic_mcc(n=4, np_cell=1)
    n, h = c.n, c.h
    sp = cfg.argon_species()
    E = cfg.argon_energy_grid(200)
    base = [((0, 2), (0, 2), 0.0, np.full_like(E, 1e-19)),
            ((0, 2), (0, 2), 11.5, cfg.lotz_like(E, cfg.S0_EXCITATION, 11.5)),
            ((0, 2), (0, 1, 0), 15.8, cfg.lotz_like(E, cfg.S0_IONIZATION, 15.8)),
            ((1, 2), (1, 2), 0.0, np.full_like(E, 5e-19))]
    dummy = base + [((0, 1), (0, 1), 0.0, np.zeros_like(E))]
    N = 60_000
    L = n[0] * h[0]
    speed = np.sqrt(2 * rng.uniform(1, 200, N) * QE / cfg.ME)
    dirs = rng.standard_normal((N, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    ve = speed[:, None] * dirs
    pos = L * rng.random((N, 3))
    dead = (rng.random(N) < 0.03).astype(np.int8)
    np_red = torch.as_tensor(np.abs(rng.standard_normal((2,) + cfg.node_shape(n))) * 1e15,
                             device=gpu)
    runs = []
    for reactions in (base, dummy):
        t = cfg.build_reaction_tables(sp, c.ngas, reactions, [(E.copy(), r[3]) for r in reactions])
        tables = P.ReactionTables(t, c.Nm, gpu)
        e = P.ParticleSet(3 * N, "cuda:0")
        e.from_host(pos[:, 0], pos[:, 1], pos[:, 2], ve[:, 0], ve[:, 1], ve[:, 2], dead=dead)
        ion = P.ParticleSet(3 * N, "cuda:0")
        ion.from_host(pos[:, 2], pos[:, 1], pos[:, 0], np.zeros(N), np.zeros(N), np.zeros(N))
        st, _, Pcoll = _run_gpu(P, [e, ion], n, h, np_red, tables, c.nb_step_collisions, c.dt,
                                _native.MCC_PHILOX, key=11, counter=4)
        runs.append((st, Pcoll, e.to_host(), ion.to_host(), e.n, ion.n))
    (sa, pa, ea, ia, nea, nia), (sb, pb, eb, ib, neb, nib) = runs
    assert sa["err_coll"][0] == 0 and sb["err_coll"][0] > 0
    for k in ("accepted", "appended", "skipped_dead"):
        assert sa[k][:2] == sb[k][:2], k
    assert sa["accepted"][0] > 0 and sa["appended"][0] > 0
    assert (nea, nia) == (neb, nib)
    for k in ("x", "y", "z", "vx", "vy", "vz", "dead"):
        np.testing.assert_array_equal(ea[k][:N], eb[k][:N], err_msg=k)
    # products land in atomic slot order: compare them as sets
    for ga, gb, m in ((ea, eb, nea), (ia, ib, nia)):
        ka = np.stack([ga[k][N:m] for k in ("x", "y", "z", "vx", "vy", "vz")], 1)
        kb = np.stack([gb[k][N:m] for k in ("x", "y", "z", "vx", "vy", "vz")], 1)
        np.testing.assert_array_equal(ka[np.lexsort(ka.T[::-1])], kb[np.lexsort(kb.T[::-1])])
    np.testing.assert_allclose(pa, pb, rtol=1e-12, atol=0)


@pytest.mark.parametrize("mode", [_native.MCC_PHILOX, _native.MCC_REPLAY_RAN2])
def test_mcc_deposited_density_matches_deposit_after(gpu, rng, mode):
    """The production collision step deposits before MCC and lets MCC add its
    products and take out its kills: the density must equal the deposit of
    the particles after MCC (the reference's order, mod_simulation.f90:425-451),
    to the summation-order tolerance of the deposit."""
    from essaitemporaire_lepic_module_amd import fields as F
    from essaitemporaire_lepic_module_amd import particles as P
    c = cfg.cfg2_pic_mcc(n=6, np_cell=1)
    n, h = c.n, c.h
    sp = cfg.argon_species()
    E = cfg.argon_energy_grid(200)
    reactions = [((0, 2), (0, 2), 0.0, np.full_like(E, 1e-19)),
                 ((0, 2), (0, 1, 0), 15.8, cfg.lotz_like(E, cfg.S0_IONIZATION, 15.8)),
                 ((0, 2), (1,), 0.0, np.full_like(E, 2e-20)),      # the electron is lost
                 ((1, 2), (1, 2), 0.0, np.full_like(E, 5e-19))]
    t = cfg.build_reaction_tables(sp, c.ngas, reactions, [(E.copy(), r[3]) for r in reactions])
    tables = P.ReactionTables(t, c.Nm, gpu)
    N = 30_000 if mode == _native.MCC_PHILOX else 4_000
    L = n[0] * h[0]
    speed = np.sqrt(2 * rng.uniform(1, 300, N) * QE / cfg.ME)
    dirs = rng.standard_normal((N, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    ve = speed[:, None] * dirs
    pos = L * rng.random((N, 3))
    ipos = L * rng.random((N, 3))
    vi = 3e4 * rng.standard_normal((N, 3))
    dead = (rng.random(N) < 0.02).astype(np.int8) * 2
    e = P.ParticleSet(3 * N + 64, "cuda:0")
    e.from_host(pos[:, 0], pos[:, 1], pos[:, 2], ve[:, 0], ve[:, 1], ve[:, 2], dead=dead)
    ion = P.ParticleSet(3 * N + 64, "cuda:0")
    ion.from_host(ipos[:, 0], ipos[:, 1], ipos[:, 2], vi[:, 0], vi[:, 1], vi[:, 2])
    bc = torch.as_tensor(cfg.grounded_box_bcnd(n), device=gpu)
    kq = F.build_kq(bc, n)
    shape = (2,) + cfg.node_shape(n)
    np_dep = torch.zeros(shape, dtype=torch.float64, device=gpu)
    for s, ps in enumerate((e, ion)):
        F.deposit_particle_set_to_np_thread(ps, n, h, kq, c.Nm[s], np_dep[s])
    np_red = torch.as_tensor(np.abs(rng.standard_normal(shape)) * 1e15, device=gpu)
    Pcoll = torch.zeros(t["ntype_total"], dtype=torch.float64, device=gpu)
    st = P.perform_collisions_step([e, ion], n, h, np_red, tables, c.nb_step_collisions, c.dt,
                                   [123456], Pcoll, mode=mode, philox_key=5, philox_counter=7,
                                   np_dep=np_dep, kq=kq)
    assert st["appended"][0] > 0 and st["appended"][1] > 0
    assert (e.to_host()["dead"] == 1).any()          # the loss channel fired
    want = torch.zeros(shape, dtype=torch.float64, device=gpu)
    for s, ps in enumerate((e, ion)):
        F.deposit_particle_set_to_np_thread(ps, n, h, kq, c.Nm[s], want[s])
    w, g = want.cpu().numpy(), np_dep.cpu().numpy()
    np.testing.assert_allclose(g, w, rtol=1e-12, atol=1e-12 * np.abs(w).max())
