"""Shared inputs for the parity tests: small grids, boundary labels, random
particle states and fields.  Everything is seeded."""
from __future__ import annotations

import numpy as np

from essaitemporaire_lepic_module_amd import config as cfg


def box(n=(8, 8, 8), side=None):
    h = (3.125e-4,) * 3 if side is None else tuple(side / d for d in n)
    return tuple(n), h


def random_particles(rng, n, h, count, margin=0.0, vscale=1e5, outside=0.0):
    """Uniform positions in [-outside, L+outside) (margin shrinks the box)."""
    L = np.array([n[0] * h[0], n[1] * h[1], n[2] * h[2]])
    lo = -outside * L + margin * L
    hi = L + outside * L - margin * L
    pos = lo + (hi - lo) * rng.random((count, 3))
    vel = vscale * rng.standard_normal((count, 3))
    return pos[:, 0], pos[:, 1], pos[:, 2], vel[:, 0], vel[:, 1], vel[:, 2]


def random_E4(rng, n, scale=1e3):
    E4 = np.zeros(cfg.node_shape(n) + (4,))
    E4[..., :3] = scale * rng.standard_normal(cfg.node_shape(n) + (3,))
    return E4


def periodic_yz_bcnd(n):
    """Walls (label 1) at x = 0 and x = L, periodic (0) planes at y and z
    boundaries, plasma elsewhere."""
    b = np.full(cfg.node_shape(n), -1, np.int32)
    b[:, :, :2] = 1
    b[:, :, n[0] + 1:] = 1
    b[:, (1, n[1] + 1), 2:n[0] + 1] = 0
    b[(1, n[2] + 1), :, 2:n[0] + 1] = 0
    b[:, 0, 2:n[0] + 1] = 0
    b[:, n[1] + 2, 2:n[0] + 1] = 0
    b[0, :, 2:n[0] + 1] = 0
    b[n[2] + 2, :, 2:n[0] + 1] = 0
    return b


def two_electrode_bcnd(n):
    """Grounded box (label 1) with an x = L face of its own label (2) and an
    internal rod of label 3."""
    b = cfg.biased_wall_bcnd(n, 2)
    cx, cy = n[0] // 2 + 1, n[1] // 2 + 1
    b[2:n[2] // 2 + 1, cy, cx] = 3
    return b


def thin_gap_bcnd(n):
    """Biased box with electrode plates one plasma node apart along z and
    y (k = 1 wall, plates at k = 3 and 5; a y plate at j = 3): the E wall
    pass reads electrode nodes its loop has already visited."""
    b = cfg.biased_wall_bcnd(n, 2)
    b[3, 3:n[1] - 1, 3:n[0]] = 3
    b[5, 4:n[1] - 2, 4:n[0] - 1] = 4
    b[6:n[2], 3, 2:n[0]] = 5
    return b
