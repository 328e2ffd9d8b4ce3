"""The combined-run configuration builders: the annular channel's labels,
the loadable-cell count the loaders fill, and the radial Br(z) map, each
against a brute-force restatement or a closed form."""
import math

import numpy as np
import pytest

from essaitemporaire_lepic_module_amd import config as cfg


def _loadable_brute(b, n):
    """Cells (i, j, k) in 1..n with a plasma node among their 8 corners,
    counted cell by cell."""
    n1, n2, n3 = n
    cnt = 0
    for k in range(1, n3 + 1):
        for j in range(1, n2 + 1):
            for i in range(1, n1 + 1):
                if (b[k:k + 2, j:j + 2, i:i + 2] == -1).any():
                    cnt += 1
    return cnt


@pytest.mark.parametrize("n", [(6, 5, 4), (10, 10, 10)])
def test_loadable_cells_grounded_box_is_every_cell(n):
    """Every cell of a grounded box touches an interior (plasma) node."""
    b = cfg.grounded_box_bcnd(n)
    assert cfg.loadable_cells(b, n) == n[0] * n[1] * n[2] == _loadable_brute(b, n)


def test_loadable_cells_matches_brute_force_on_the_channel():
    n, h = (16, 16, 12), (3.125e-4,) * 3
    b = cfg.annular_channel_bcnd(n, h)
    got = cfg.loadable_cells(b, n)
    assert got == _loadable_brute(b, n)
    assert 0 < got < n[0] * n[1] * n[2]


def test_annular_channel_labels():
    n, h = (24, 24, 16), (3.125e-4,) * 3
    b = cfg.annular_channel_bcnd(n, h, r_in=0.3, r_out=0.45, anode_label=2)
    assert b.shape == cfg.node_shape(n)
    L = n[0] * h[0]
    for k in range(b.shape[0]):
        for j in range(b.shape[1]):
            for i in range(b.shape[2]):
                r = math.hypot((i - 1) * h[0] - L / 2, (j - 1) * h[1] - L / 2)
                chan = 0.3 * L <= r <= 0.45 * L
                interior = 2 <= i <= n[0] and 2 <= j <= n[1] and 2 <= k <= n[2]
                if chan and k <= 1:
                    want = 2                          # the anode face at z = 0
                elif chan and interior:
                    want = -1
                else:
                    want = 1
                assert b[k, j, i] == want, (i, j, k)


def test_radial_map_closed_form():
    """B = Br(z) r_hat, Br(z) = B0 exp(-(z - 0.7 L)^2 / (2 (0.1 L)^2)): zero on
    the axis, |B| = Br(z) off it, no z component, peak B0 at z = 0.7 L."""
    n, h = (20, 20, 20), (3.125e-4,) * 3
    B4 = cfg.radial_bfield_map(n, h, B0=0.02, z_peak=0.7, width=0.1)
    L = n[2] * h[2]
    k_peak = 1 + round(0.7 * n[2])                    # node z = 0.7 L
    i0 = 1 + n[0] // 2                                # node x = L / 2 (the axis)
    np.testing.assert_array_equal(B4[:, i0, i0, :], 0.0)
    np.testing.assert_array_equal(B4[..., 2], 0.0)
    np.testing.assert_allclose(B4[..., 3], np.hypot(B4[..., 0], B4[..., 1]), rtol=1e-15)
    for k in (1, 5, k_peak, 19):
        z = (k - 1) * h[2]
        br = 0.02 * math.exp(-(z - 0.7 * L) ** 2 / (2 * (0.1 * L) ** 2))
        off = B4[k][B4[k, ..., 3] > 0, 3]
        np.testing.assert_allclose(off, br, rtol=1e-14)
    assert B4[k_peak, i0, i0 + 3, 3] == pytest.approx(0.02, rel=1e-14)
    # r_hat points away from the axis: +x on the +x side, -y on the -y side
    assert B4[k_peak, i0, i0 + 3, 0] > 0 and B4[k_peak, i0 - 3, i0, 1] < 0


def test_cfg5_case():
    c = cfg.cfg5_combined(n=16, np_cell=4, steps=10)
    assert c.n == (16, 16, 16) and c.ngrid == 2
    assert c.electrode_V == {1: 0.0, 2: 300.0}
    assert c.B_map.shape == cfg.node_shape(c.n) + (4,)
    assert c.chemistry is not None and c.nb_step_collisions == 10 and c.nb_step_sort == 20
    assert c.eps == 1e-10
    assert c.flux_plane == (2, 2, -1.0, 1.0)
    # the channel holds the 0.45^2 - 0.3^2 of pi L^2 annulus and its edges
    frac = cfg.loadable_cells(c.bcnd, c.n) / 16 ** 3
    assert math.pi * (0.45 ** 2 - 0.3 ** 2) < frac < 0.6
