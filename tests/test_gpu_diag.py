"""Plane moments on the device (lepic_plane_moments) against the oracle's
restatement of mod_planeMoments.f90, and the driver's diagnostics attached
to a Simulation (cadences, .mco files, the printed summary).

The plane sums go through FP64 atomics, so counts, means and temperatures
are held to 1e-11 relative to each array's largest value; everything else
(which nodes receive weight, the untouched nodes) is exact."""
import os

import numpy as np
import pytest
import torch

from essaitemporaire_lepic_module_amd import config as cfg
from essaitemporaire_lepic_module_amd import diagnostics as DG

pytestmark = pytest.mark.gpu

RTOL = 1e-11


def _random_state(rng, n, h, m):
    L = np.array(n) * np.array(h)
    pos = [rng.uniform(-0.05 * L[d], 1.05 * L[d], m) for d in range(3)]
    vel = [rng.normal(0.0, 3e5, m) + 2e4 * d for d in range(3)]
    dead = (rng.random(m) < 0.05).astype(np.int8)
    return pos, vel, dead


def test_plane_moments_match_oracle(gpu, oracle):
    from essaitemporaire_lepic_module_amd import particles as P
    n, h = (16, 12, 10), (1.0e-3, 1.2e-3, 0.9e-3)
    rng = np.random.default_rng(21)
    m = 300000
    pos, vel, dead = _random_state(rng, n, h, m)
    op = oracle.Particles(m)
    op.set(*pos, *vel)
    op.dead[:m] = dead
    dp = P.ParticleSet(m + 16, "cuda:0")
    dp.from_host(*pos, *vel, dead=dead)
    shapes = DG.plane_shapes(n)
    host = [rng.normal(0, 1, s + (5,)) for s in shapes]        # data already holding values
    dev = [torch.as_tensor(a.copy(), device="cuda:0") for a in host]
    planes = (9, 7, 5)
    for call, dt in enumerate((2e-10, 5e-10)):
        oracle.plane_moments(n, h, op, dt, cfg.ME, *planes, *host)
        DG.compute_particle_plane_moments_species(dp, n, h, dt, cfg.ME, *planes, *dev)
        torch.cuda.synchronize()
        for k, (a, b) in enumerate(zip(host, dev)):
            g = b.cpu().numpy()
            for comp in range(5):
                ref = a[..., comp]
                np.testing.assert_allclose(g[..., comp], ref, rtol=0,
                                           atol=RTOL * max(np.abs(ref).max(), 1e-300),
                                           err_msg=f"call {call} plane {k} comp {comp}")


def test_plane_moments_empty_and_bad_plane(gpu):
    from essaitemporaire_lepic_module_amd import _native
    from essaitemporaire_lepic_module_amd import particles as P
    n, h = (8, 8, 8), (1e-3,) * 3
    dp = P.ParticleSet(16, "cuda:0")
    dev = [torch.full(s + (5,), 3.0, dtype=torch.float64, device="cuda:0")
           for s in DG.plane_shapes(n)]
    DG.compute_particle_plane_moments_species(dp, n, h, 1e-9, cfg.ME, 5, 5, 5, *dev)
    torch.cuda.synchronize()
    for d in dev:
        assert (d == 3.0).all()          # no particle: nothing overwritten
    with pytest.raises(_native.LepicError):
        DG.compute_particle_plane_moments_species(dp, n, h, 1e-9, cfg.ME, 0, 5, 5, *dev)


def test_simulation_diagnostics(gpu, tmp_path):
    """Moments and plane sums on steps 1, 6, 11 (nsav = 5), files and the
    reset after steps 1, 6, 11, the summary before them."""
    import io
    from essaitemporaire_lepic_module_amd.simulation import Simulation
    case = cfg.cfg1_plain_box(n=8, np_cell=6, ng=3)
    sim = Simulation(case, "cuda:0", fused=True, lazy=True)
    d = DG.Diagnostics(sim, nsav=5, out_dir=str(tmp_path), names=["e", "Ar+"])
    sim.diagnostics = d
    d.write_initial_diagnostics()
    out = io.StringIO()
    sim.run(12, out=out)
    torch.cuda.synchronize()
    text = out.getvalue()
    assert text.count(" TIME STEP ") == 3 and " npart e = " in text and " Pwall (W)  = " in text
    for it in (1, 6, 11):
        for stem in ("n1", "n2", "T1", "T2", "phi", "Ex", "Ey", "Ez"):
            for pl in ("xy", "xz", "yz"):
                assert os.path.exists(tmp_path / f"it{it}_{stem}_{pl}.mco"), (it, stem, pl)
    for stem in ("n1", "n2", "phi1", "Ex", "Ey", "Ez"):
        assert os.path.exists(tmp_path / f"{stem}_xy.mco")
    # after step 11 the sums were written and reset
    assert d.cnt_avg == 0 and float(d.np_avg_xy.abs().sum()) == 0.0
    # the header and row count of one file: (n+1) rows of n+1 values
    lines = (tmp_path / "it6_phi_xz.mco").read_text().splitlines()
    assert lines[0].split() == ["9", "9"] and len(lines) == 10
    assert all(len(r.split()) == 9 for r in lines[1:])
    # step 1 takes the moments and the density plane of the step's np before
    # the push (and, since mod(1, nsav) == 1, resets them after the deposit)
    sim2 = Simulation(case, "cuda:0", fused=True, lazy=True)
    d2 = DG.Diagnostics(sim2, nsav=50)
    sim2.diagnostics = d2
    seen = {}

    def hook(istep, orig=d2.before_push):
        seen["np"] = sim2.np_red[:, d2.iz_plane].clone()
        orig(istep)
        seen["cnt"] = d2.cnt_avg
        seen["avg"] = d2.np_avg_xy.clone()
        seen["w"] = float(d2.data_pavg_xy[..., DG.NP_AVG].sum())
    d2.before_push = hook
    sim2.advance_one_step(1)
    torch.cuda.synchronize()
    assert seen["cnt"] == 1 and torch.equal(seen["avg"], seen["np"]) and seen["w"] > 0.0
    assert d2.cnt_avg == 0


def test_backup_and_restart(gpu, tmp_path):
    """A production run's backup restarts into a Simulation holding the same
    live particles in the same order and the same deposited density."""
    from essaitemporaire_lepic_module_amd import restart as R
    from essaitemporaire_lepic_module_amd.simulation import Simulation
    case = cfg.cfg1_plain_box(n=8, np_cell=6, ng=3)
    sim = Simulation(case, "cuda:0", fused=True, lazy=True)
    sim.run(3)
    torch.cuda.synchronize()
    path = R.save_simulation(sim, str(tmp_path / "DATA.BAK"), time=3 * case.dt)
    assert os.path.basename(path) == "particles0.bak"
    sim2, t = R.restart_simulation(case, path, "cuda:0", fused=True, lazy=True)
    assert t == 3 * case.dt
    for a, b in zip(sim.parts, sim2.parts):
        ha, hb = a.to_host(), b.to_host()
        live = ha["dead"] == 0
        assert int(live.sum()) == b.n
        for k in ("x", "y", "z", "vx", "vy", "vz"):
            np.testing.assert_array_equal(ha[k][live], hb[k])
    np.testing.assert_allclose(sim2.np_thread.cpu().numpy(), sim.np_thread.cpu().numpy(),
                               rtol=0, atol=1e-12 * float(sim.np_thread.abs().max()))
    sim2.run(2, start=4)
    torch.cuda.synchronize()
    assert sim2.particle_count() > 0
