"""Golden vectors of the reference's own Fortran routines
(tests/golden/reference_vectors.npz, written by tests/golden/make_golden.py
from the reference's modules built in oracle/_ref).

The oracle must reproduce them bit for bit on the CPU, and so must the HIP
entry points on the GPU, except the deposition, whose summation order on the
GPU is that of its atomics (1e-12 relative)."""
import os

import numpy as np
import pytest
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "reference_vectors.npz"))
H3 = (3.125e-4,) * 3
KEYS = ("x", "y", "z", "vx", "vy", "vz")


def _in(prefix):
    return [GOLD[f"{prefix}_in_{k}"] for k in KEYS]


def _dims(shape):
    """(n1, n2, n3) of a node array shaped (n3+3, n2+3, n1+3)."""
    return (shape[2] - 3, shape[1] - 3, shape[0] - 3)


# ----------------------------------------------------------------- oracle (CPU)
def test_oracle_ran2_golden(oracle):
    draws, seed = oracle.ran2(123456, len(GOLD["ran2_draws"]))
    np.testing.assert_array_equal(draws, GOLD["ran2_draws"])
    assert seed == int(GOLD["ran2_seed_after"][0])


def test_abi_ran2_golden():
    from essaitemporaire_lepic_module_amd import particles as P
    s = [123456]
    got = [P.ran2(s) for _ in range(len(GOLD["ran2_draws"]))]
    np.testing.assert_array_equal(got, GOLD["ran2_draws"])
    assert s[0] == int(GOLD["ran2_seed_after"][0])


def test_oracle_sort_golden(oracle):
    n = tuple(int(v) for v in GOLD["sort_n"])
    p = oracle.Particles(len(GOLD["sort_in_x"])).set(*_in("sort"), dead=GOLD["sort_in_dead"])
    cc, cs = oracle.sort_particles(n, H3, p)
    a = p.arrays()
    for k in KEYS + ("dead",):
        np.testing.assert_array_equal(a[k], GOLD[f"sort_out_{k}"])
    np.testing.assert_array_equal(a["cell"] + 1, GOLD["sort_out_cell1"])
    np.testing.assert_array_equal(cc, GOLD["sort_out_count"])
    np.testing.assert_array_equal(cs + 1, GOLD["sort_out_start1"])


def test_oracle_move_golden(oracle):
    n = tuple(int(v) for v in GOLD["move_n"])
    q, m, dt = GOLD["move_qmdt"]
    p = oracle.Particles(len(GOLD["move_in_x"])).set(*_in("move"))
    oracle.move_electrostatic(n, H3, p, GOLD["move_E4"], q, m, dt)
    a = p.arrays()
    for k in KEYS:
        np.testing.assert_array_equal(a[k], GOLD[f"move_out_{k}"])


@pytest.mark.parametrize("tag", ["bc0", "bc1"])
def test_oracle_apply_bc_golden(oracle, tag):
    bc = GOLD[f"{tag}_bcnd"]
    n = _dims(bc.shape)
    pbc, nmn, ptype, tagn, ngrid = (int(v) for v in GOLD[f"{tag}_params"])
    qmacro, mass, Nm = GOLD[f"{tag}_qmN"]
    p = oracle.Particles(len(GOLD[f"{tag}_in_x"])).set(*_in(tag), dead=GOLD[f"{tag}_in_dead"])
    lost, pmac, Ploss = oracle.apply_bc(n, H3, p, bc, pbc, nmn, ptype, tagn, qmacro, mass, Nm, ngrid)
    a = p.arrays()
    for k in KEYS + ("dead",):
        np.testing.assert_array_equal(a[k], GOLD[f"{tag}_out_{k}"])
    np.testing.assert_array_equal(pmac.reshape(ngrid + 1, 2)[1:], GOLD[f"{tag}_out_p_mac_slots"])
    assert Ploss == GOLD[f"{tag}_out_P_loss"][0]


def test_oracle_deposit_golden(oracle):
    bc = GOLD["dep_bcnd"]
    n = _dims(bc.shape)
    kq = oracle.build_kq(n, bc)
    np.testing.assert_array_equal(kq, GOLD["dep_kq"])
    p = oracle.Particles(len(GOLD["dep_in_x"])).set(
        GOLD["dep_in_x"], GOLD["dep_in_y"], GOLD["dep_in_z"], *(np.zeros_like(GOLD["dep_in_x"]),) * 3,
        dead=GOLD["dep_in_dead"])
    np.testing.assert_array_equal(oracle.deposit(n, H3, p, kq, GOLD["dep_Nm"][0]), GOLD["dep_out_np"])


@pytest.mark.parametrize("tag", ["e0", "e1", "e2"])
def test_oracle_efield_golden(oracle, tag):
    bc = GOLD[f"{tag}_bcnd"]
    E4 = oracle.efield(_dims(bc.shape), H3, GOLD[f"{tag}_phi"], bc)
    np.testing.assert_array_equal(E4[..., :3], GOLD[f"{tag}_out_E"])


# -------------------------------------------------------------------- HIP (GPU)
def _gpu_set(P, arrays, dead=None):
    ps = P.ParticleSet(len(arrays[0]) + 16, "cuda:0")
    ps.from_host(*arrays, dead=dead)
    return ps


@pytest.mark.gpu
def test_gpu_sort_golden(gpu):
    from essaitemporaire_lepic_module_amd import particles as P
    n = tuple(int(v) for v in GOLD["sort_n"])
    ps = _gpu_set(P, _in("sort"), dead=GOLD["sort_in_dead"])
    P.sort_particles_by_cell(ps, n, H3, cell_tables=True)
    a = ps.to_host()
    for k in KEYS + ("dead",):
        np.testing.assert_array_equal(a[k], GOLD[f"sort_out_{k}"])
    np.testing.assert_array_equal(a["cell"] + 1, GOLD["sort_out_cell1"])
    np.testing.assert_array_equal(ps.cell_count.cpu().numpy(), GOLD["sort_out_count"])
    np.testing.assert_array_equal(ps.cell_start.cpu().numpy() + 1, GOLD["sort_out_start1"])


@pytest.mark.gpu
def test_gpu_move_golden(gpu):
    from essaitemporaire_lepic_module_amd import particles as P
    n = tuple(int(v) for v in GOLD["move_n"])
    q, m, dt = (float(v) for v in GOLD["move_qmdt"])
    ps = _gpu_set(P, _in("move"))
    P.move_particles_electrostatic(ps, n, H3, torch.as_tensor(GOLD["move_E4"], device=gpu), q, m, dt)
    a = ps.to_host()
    for k in KEYS:
        np.testing.assert_array_equal(a[k], GOLD[f"move_out_{k}"])


@pytest.mark.gpu
@pytest.mark.parametrize("tag", ["bc0", "bc1"])
def test_gpu_apply_bc_golden(gpu, tag):
    from essaitemporaire_lepic_module_amd import particles as P
    bc = GOLD[f"{tag}_bcnd"]
    n = _dims(bc.shape)
    pbc, nmn, ptype, tagn, ngrid = (int(v) for v in GOLD[f"{tag}_params"])
    qmacro, mass, Nm = (float(v) for v in GOLD[f"{tag}_qmN"])
    ps = _gpu_set(P, _in(tag), dead=GOLD[f"{tag}_in_dead"])
    p_mac = torch.zeros((ngrid + 1, 2), dtype=torch.float64, device=gpu)
    P_loss = torch.zeros(1, dtype=torch.float64, device=gpu)
    P.apply_particle_bc_legacy(ps, n, H3, torch.as_tensor(bc, device=gpu), pbc, nmn, ptype, tagn, 0,
                               qmacro, p_mac, mass, P_loss, Nm)
    a = ps.to_host()
    for k in KEYS + ("dead",):
        np.testing.assert_array_equal(a[k], GOLD[f"{tag}_out_{k}"])
    np.testing.assert_allclose(p_mac.cpu().numpy()[1:], GOLD[f"{tag}_out_p_mac_slots"], rtol=1e-12)
    np.testing.assert_allclose(P_loss.cpu().numpy(), GOLD[f"{tag}_out_P_loss"], rtol=1e-12)


@pytest.mark.gpu
def test_gpu_deposit_golden(gpu):
    from essaitemporaire_lepic_module_amd import fields as F
    from essaitemporaire_lepic_module_amd import particles as P
    bc = GOLD["dep_bcnd"]
    n = _dims(bc.shape)
    bcg = torch.as_tensor(bc, device=gpu)
    kq = F.build_kq(bcg, n)
    np.testing.assert_array_equal(kq.cpu().numpy(), GOLD["dep_kq"])
    z = np.zeros_like(GOLD["dep_in_x"])
    ps = _gpu_set(P, [GOLD["dep_in_x"], GOLD["dep_in_y"], GOLD["dep_in_z"], z, z, z],
                  dead=GOLD["dep_in_dead"])
    npl = torch.zeros(bc.shape, dtype=torch.float64, device=gpu)
    F.deposit_particle_set_to_np_thread(ps, n, H3, kq, float(GOLD["dep_Nm"][0]), npl)
    want = GOLD["dep_out_np"]
    np.testing.assert_allclose(npl.cpu().numpy(), want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


@pytest.mark.gpu
@pytest.mark.parametrize("tag", ["e0", "e1", "e2"])
def test_gpu_efield_golden(gpu, tag):
    from essaitemporaire_lepic_module_amd import fields as F
    bc = GOLD[f"{tag}_bcnd"]
    n = _dims(bc.shape)
    E4 = torch.zeros(bc.shape + (4,), dtype=torch.float64, device=gpu)
    F.calc_Efield_modular(n, H3, torch.as_tensor(GOLD[f"{tag}_phi"], device=gpu), E4,
                          torch.as_tensor(bc, device=gpu))
    np.testing.assert_array_equal(E4.cpu().numpy()[..., :3], GOLD[f"{tag}_out_E"])
