"""The mixed-precision Poisson solve (FP32 red-black smoothing inside FP64
iterative refinement) against the oracle's FP64 multigrid (pinned by the
direct solve and the eigenmode tests of test_oracle.py): both converged to
the same L1 / alpha tolerance, phi agrees to the tolerance the residual
bound allows, and E built from it agrees as well."""
import numpy as np
import pytest
import torch

from essaitemporaire_lepic_module_amd import config as cfg
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _solve_both(oracle, gpu, n, h, bc, phi0, rhs, ng, eps, flag_pbc=0, flag_nmn=0, **kw):
    from essaitemporaire_lepic_module_amd import fields as F
    phi_o = phi0.copy()
    st_o, info_o = oracle.pdesolver(n, h, phi_o, rhs.copy(), bc, flag_pbc=flag_pbc,
                                    flag_nmn=flag_nmn, eps=eps, omega=1.4, ng=ng)
    assert st_o == 0
    phi_g = torch.as_tensor(phi0.copy(), device=gpu)
    info = F.solve_poisson_mixed(phi_g, torch.as_tensor(bc, device=gpu),
                                 torch.as_tensor(rhs, device=gpu), h, n, 20000, eps, 1.4, ng,
                                 flag_pbc, flag_nmn, **kw)
    return phi_o, phi_g.cpu().numpy(), info_o, info


@pytest.mark.parametrize("n,ng", [(32, 5), (64, 6)])
def test_mixed_matches_fp64_faces_and_blobs(gpu, oracle, n, ng):
    """cfg3's problem at 32^3 and 64^3 (six biased faces, Gaussian blobs):
    the mixed solve reaches the FP64 solve's tolerance (1e-10), phi within
    1e-8 of max |phi| of the oracle's FP64 solution, E within 1e-6 of
    max |E|, and electrode nodes untouched."""
    n3, h, bc, phi0, rhs = cfg.cfg3_poisson(n=n, seed=5, nblobs=16)
    phi_o, phi_g, info_o, info = _solve_both(oracle, gpu, n3, h, bc, phi0, rhs, ng, 1e-10)
    assert info["res"] < 1e-10 and info["cycles"] >= 1
    scale = np.abs(phi_o).max()
    assert np.abs(phi_g - phi_o).max() <= 1e-8 * scale
    np.testing.assert_array_equal(phi_g[bc > 0], phi0[bc > 0])
    E_o = oracle.efield(n3, h, phi_o, bc)
    E_g = oracle.efield(n3, h, phi_g, bc)
    assert np.abs(E_g - E_o).max() <= 1e-6 * np.abs(E_o).max()


def test_mixed_periodic_and_neumann(gpu, oracle, rng):
    """Periodic y/z with walls in x, and a Neumann x = 0 face: the same
    agreement, ghosts included."""
    n = (16, 16, 16)
    h = (3.125e-4,) * 3
    bc = H.periodic_yz_bcnd(n)
    phi0 = np.zeros(bc.shape)
    phi0[bc == 1] = 3.0
    rhs = np.where(bc == -1, 1e-6 * rng.standard_normal(bc.shape), 0.0)
    phi_o, phi_g, _, info = _solve_both(oracle, gpu, n, h, bc, phi0, rhs, 4, 1e-10, flag_pbc=1)
    assert info["res"] < 1e-10
    assert np.abs(phi_g - phi_o).max() <= 1e-8 * np.abs(phi_o).max()
    bcn = H.two_electrode_bcnd(n)
    bcn[2:n[2] + 1, 2:n[1] + 1, 0:2] = -2
    phi0 = np.zeros(bcn.shape)
    phi0[bcn == 2] = -5.0
    rhs = np.where(bcn == -1, 1e-6 * rng.standard_normal(bcn.shape), 0.0)
    phi_o, phi_g, _, info = _solve_both(oracle, gpu, n, h, bcn, phi0, rhs, 4, 1e-10, flag_nmn=1)
    assert info["res"] < 1e-10
    # the ghost plane i = 0 behind the Neumann face is read by no row (the
    # mirror replaces it) and only collects prolonged corrections, which the
    # two solvers sum differently: compare the physical nodes
    phys = (slice(1, n[2] + 2), slice(1, n[1] + 2), slice(1, n[0] + 2))
    assert np.abs(phi_g[phys] - phi_o[phys]).max() <= 1e-8 * np.abs(phi_o).max()


def test_mixed_warm_start_and_graph(gpu, oracle):
    """A warm start that already meets the tolerance takes no refinement
    step; replaying the captured refinement step (graph) and launching it
    eagerly give the same phi bit for bit."""
    from essaitemporaire_lepic_module_amd import fields as F
    n3, h, bc, phi0, rhs = cfg.cfg3_poisson(n=32, seed=9, nblobs=8)
    bcg, rg = torch.as_tensor(bc, device=gpu), torch.as_tensor(rhs, device=gpu)
    outs = []
    for graph in (True, False):
        phi = torch.as_tensor(phi0.copy(), device=gpu)
        F.solve_poisson_mixed(phi, bcg, rg, h, n3, 20000, 1e-10, 1.4, 5, 0, 0, graph=graph)
        outs.append(phi.cpu().numpy())
        info = F.solve_poisson_mixed(phi, bcg, rg, h, n3, 20000, 1e-10, 1.4, 5, 0, 0, graph=graph)
        assert info["cycles"] == 0 and info["res"] < 1e-10
    np.testing.assert_array_equal(outs[0], outs[1])


def _pcg_vs_oracle(oracle, gpu, n, h, bc, phi0, rhs, ng, eps, **kw):
    from essaitemporaire_lepic_module_amd import fields as F
    phi_o = phi0.copy()
    st_o, _ = oracle.pdesolver(n, h, phi_o, rhs.copy(), bc, eps=eps, omega=1.4, ng=ng)
    assert st_o == 0
    phi_g = torch.as_tensor(phi0.copy(), device=gpu)
    info = F.solve_poisson_pcg(phi_g, torch.as_tensor(bc, device=gpu),
                               torch.as_tensor(rhs, device=gpu), h, n, 20000, eps, 1.4, ng, **kw)
    return phi_o, phi_g.cpu().numpy(), info


@pytest.mark.parametrize("n,ng,inner,fp64", [(32, 5, 1, False), (64, 6, 1, False),
                                              (64, 6, 2, False), (32, 5, 1, True),
                                              (64, 6, 1, True)])
def test_pcg_matches_fp64_faces_and_blobs(gpu, oracle, n, ng, inner, fp64):
    """cfg3's problem by the preconditioned conjugate gradients (FP32 or FP64
    V-cycles as the preconditioner): the same tolerance, phi and E agreement
    and untouched electrodes as the mixed solve."""
    n3, h, bc, phi0, rhs = cfg.cfg3_poisson(n=n, seed=5, nblobs=16)
    phi_o, phi_g, info = _pcg_vs_oracle(oracle, gpu, n3, h, bc, phi0, rhs, ng, 1e-10,
                                        inner_cycles=inner, precond_fp64=fp64)
    assert info["res"] < 1e-10 and info["cycles"] >= 1
    scale = np.abs(phi_o).max()
    assert np.abs(phi_g - phi_o).max() <= 1e-8 * scale
    np.testing.assert_array_equal(phi_g[bc > 0], phi0[bc > 0])
    E_o = oracle.efield(n3, h, phi_o, bc)
    E_g = oracle.efield(n3, h, phi_g, bc)
    assert np.abs(E_g - E_o).max() <= 1e-6 * np.abs(E_o).max()


@pytest.mark.parametrize("n,ng,fp64", [(32, 5, False), (64, 6, False), (64, 6, True)])
def test_pcg_annular_channel(gpu, oracle, rng, n, ng, fp64):
    """The combined run's channel (curved inner and outer electrodes, the
    anode at 300 V on z = 0) with a random charge density in the plasma:
    the conjugate-gradient solve against the oracle's FP64 multigrid, in
    fewer FP32 V-cycles than the refinement solve."""
    from essaitemporaire_lepic_module_amd import fields as F
    c = cfg.cfg5_combined(n=n, np_cell=1)
    bc, h, n3 = c.bcnd, c.h, c.n
    phi0 = c.phi0()
    rhs = np.where(bc == -1, 1e-4 * rng.standard_normal(bc.shape), 0.0)
    phi_o, phi_g, info = _pcg_vs_oracle(oracle, gpu, n3, h, bc, phi0, rhs, ng, 1e-10,
                                        precond_fp64=fp64)
    assert info["res"] < 1e-10
    assert np.abs(phi_g - phi_o).max() <= 1e-8 * np.abs(phi_o).max()
    np.testing.assert_array_equal(phi_g[bc > 0], phi0[bc > 0])
    phi_m = torch.as_tensor(phi0.copy(), device=gpu)
    mix = F.solve_poisson_mixed(phi_m, torch.as_tensor(bc, device=gpu),
                                torch.as_tensor(rhs, device=gpu), h, n3, 20000, 1e-10, 1.4, ng,
                                0, 0)
    print(f"pcg iterations {info['cycles']}, mixed refinement steps {mix['cycles']} x 4")
    assert info["cycles"] < 4 * mix["cycles"]


def test_pcg_warm_start_graph_and_refusals(gpu):
    """A warm start that meets the tolerance takes no iteration; the captured
    iteration and the eager one give the same phi bit for bit over repeated
    solves; periodic or Neumann walls and labels 0 or -2 are refused."""
    from essaitemporaire_lepic_module_amd import _native
    from essaitemporaire_lepic_module_amd import fields as F
    n3, h, bc, phi0, rhs = cfg.cfg3_poisson(n=32, seed=9, nblobs=8)
    bcg, rg = torch.as_tensor(bc, device=gpu), torch.as_tensor(rhs, device=gpu)
    for fp64 in (False, True):
        outs = []
        for graph in (True, False):
            phi = torch.as_tensor(phi0.copy(), device=gpu)
            for rep in range(3):
                info = F.solve_poisson_pcg(phi, bcg, rg, h, n3, 200,
                                           1e-10 if rep == 0 else 1e-11, 1.4, 5, graph=graph,
                                           precond_fp64=fp64)
                assert info["res"] < (1e-10 if rep == 0 else 1e-11)
                torch.cuda.synchronize()
            outs.append(phi.cpu().numpy())
            info = F.solve_poisson_pcg(phi, bcg, rg, h, n3, 20000, 1e-10, 1.4, 5, graph=graph,
                                       precond_fp64=fp64)
            assert info["cycles"] == 0 and info["res"] < 1e-10
        np.testing.assert_array_equal(outs[0], outs[1])
    with pytest.raises(ValueError):
        F.solve_poisson_pcg(phi, bcg, rg, h, n3, 20000, 1e-10, 1.4, 5, flag_pbc=1)
    bad = bc.copy()
    bad[5, 5, 5] = -2
    with pytest.raises(_native.LepicError):
        F.solve_poisson_pcg(phi, torch.as_tensor(bad, device=gpu), rg, h, n3, 20000, 1e-10,
                            1.4, 5)
