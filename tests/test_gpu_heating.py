"""Electron heating on the device (lepic_electron_heating) against the
oracle.  Replay mode: the same ran2 chain (seed exact), positions and
window counts exact, velocities to 1e-14 relative (log, sqrt, sin and cos
of the two math libraries differ in the last bit), the window energy and
the absorbed energy to 1e-12 (summation order).  Philox mode: the law
(acceptance fraction, re-drawn temperature) and the energy bookkeeping."""
import numpy as np
import pytest
import torch

from essaitemporaire_lepic_module_amd import _native
from essaitemporaire_lepic_module_amd import config as cfg

pytestmark = pytest.mark.gpu

N, H = (16, 8, 8), (1e-3, 1e-3, 1e-3)


def _state(rng, m):
    L = np.array(N) * np.array(H)
    pos = [rng.uniform(0, L[d], m) for d in range(3)]
    vel = [rng.normal(0, 8e5, m) for _ in range(3)]
    return pos, vel


def _hp(P, nudt, **kw):
    d = dict(nudt=nudt, Pabs=0.2, nu_h=1e8, xl_pow=4.2e-3, xr_pow=11.7e-3, ymax=8e-3, zmax=8e-3,
             Nm=1e5, mass=cfg.ME, vt0=1e6, flag_circxh=0, flag_ahp=0, R_ahp_moments=0.0,
             R_ahp=0.0)
    d.update(kw)
    return d, P.heating_params(**d)


@pytest.mark.parametrize("circle", [False, True])
def test_heating_replay_matches_oracle(gpu, oracle, circle):
    """MCC-killed electrons (dead = 1) are still in the reference's arrays
    when heating runs and take part; holes (dead = 2) are not there."""
    from essaitemporaire_lepic_module_amd import particles as P
    rng = np.random.default_rng(7)
    m = 40000
    pos, vel = _state(rng, m)
    dead = np.zeros(m, np.int8)
    dead[rng.random(m) < 0.03] = 1
    dead[rng.random(m) < 0.05] = 2
    keep = dead != 2
    kw = dict(flag_circxh=1, flag_ahp=int(circle), R_ahp_moments=2.5e-3, R_ahp=3e-3) if circle else {}
    d, hp = _hp(P, 0.35, **kw)
    op = oracle.Particles(int(keep.sum()))
    op.set(*(a[keep] for a in pos), *(a[keep] for a in vel))
    op.dead[:op.n] = dead[keep]
    seed_o, P_o, (Nh_o, Ek_o, vt_o) = oracle.electron_heating(N, H, op, d, 4242, 1.5)
    dp = P.ParticleSet(m + 16, "cuda:0")
    dp.from_host(*pos, *vel, dead=dead)
    P_heat = torch.tensor([1.5], dtype=torch.float64, device="cuda:0")
    seed = [4242]
    st = P.apply_electron_heating(dp, N, H, hp, seed, P_heat, _native.MCC_REPLAY_RAN2)
    torch.cuda.synchronize()
    assert seed[0] == seed_o and st["Nh"] == Nh_o
    assert st["sum_Ek"] == pytest.approx(Ek_o, rel=1e-12)
    assert st["vt"] == pytest.approx(vt_o, rel=1e-14)
    assert float(P_heat[0]) == pytest.approx(P_o, rel=1e-12)
    a, b = dp.to_host(), op.arrays()
    for k in ("x", "y", "z"):
        np.testing.assert_array_equal(a[k][keep], b[k])
    for k in ("vx", "vy", "vz"):
        np.testing.assert_allclose(a[k][keep], b[k], rtol=1e-14, atol=1e-14 * np.abs(b[k]).max())
        np.testing.assert_array_equal(a[k][~keep], vel["xyz".index(k[1])][~keep])


def test_heating_philox_law(gpu):
    from essaitemporaire_lepic_module_amd import particles as P
    rng = np.random.default_rng(8)
    m = 400000
    pos, vel = _state(rng, m)
    dead = (rng.random(m) < 0.02).astype(np.int8) * 2
    d, hp = _hp(P, 0.4)
    dp = P.ParticleSet(m + 16, "cuda:0")
    dp.from_host(*pos, *vel, dead=dead)
    P_heat = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    st = P.apply_electron_heating(dp, N, H, hp, None, P_heat, _native.MCC_PHILOX,
                                  philox_key=11, philox_counter=4)
    torch.cuda.synchronize()
    a = dp.to_host()
    ix = (pos[0] / H[0]).astype(np.int64) + 1
    win = (ix >= int(d["xl_pow"] / H[0]) + 1) & (ix <= int(d["xr_pow"] / H[0]) + 1) & (dead == 0)
    changed = a["vx"] != vel[0]
    assert not changed[~win].any()
    frac = changed[win].mean()
    assert frac == pytest.approx(0.4, abs=0.01)
    vt = st["vt"]
    for k in ("vx", "vy", "vz"):
        v = a[k][changed]
        assert np.mean(v * v) == pytest.approx(vt * vt / 2, rel=0.02)
    v2a = a["vx"] ** 2 + a["vy"] ** 2 + a["vz"] ** 2
    v2b = vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2
    dE = np.sum(0.5 * d["Nm"] * cfg.ME * (v2a - v2b)[changed])
    assert float(P_heat[0]) == pytest.approx(dE, rel=1e-10)
    # a different counter draws differently
    dp2 = P.ParticleSet(m + 16, "cuda:0")
    dp2.from_host(*pos, *vel, dead=dead)
    P.apply_electron_heating(dp2, N, H, hp, None, P_heat, _native.MCC_PHILOX, philox_key=11,
                             philox_counter=8, stats=False)
    torch.cuda.synchronize()
    assert not np.array_equal(dp2.to_host()["vx"], a["vx"])
